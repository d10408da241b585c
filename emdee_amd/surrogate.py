"""Learned stand-in for the k-space electrostatics of rigid water
(EmDeeSurrogate_* and EmDeeGPU_attach_surrogate of include/emdee_gpu.h).

Two model layers describe the same water (reference mechanism for multiple
time steps, EmDeeCode.f90:273-305, 486-520, 929-946):
  layer 1: the exact model, LJ + coul_long + kspace_ewald;
  layer 2: LJ + coul_damped(alpha), i.e. coul_long's real-space part alone.
Their per-atom energy difference is the k-space (reciprocal + self) energy
of each atom.  The model reproduces it as an exact reciprocal sum over the
longest wavelengths (|k| <= kcut, when set) plus a network on descriptors of
the atom's 6 A environment trained on the remainder; attached to layer 2 it
restores an approximation of the exact model at the cost of the real-space
layer.  Exact layer-1 energies judge every Metropolis acceptance in `hmc`.
"""
import ctypes
import math
import time

import numpy as np

from .water_box import CHARGE, EPS_OO, KB, KCOUL, MASS, SIG_OO, water_box

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


def two_layer_water(lib, nmol, seed, rc=10.0, skin=2.0, accuracy=math.exp(-9.0) / 9.0, alpha=0.3, mode=0,
                    short="damped", short_alpha=None):
    """short: layer-2 Coulomb model, "damped" (coul_damped(alpha): coul_long's
    real-space part) or "dsf" (shifted-force coul_damped(short_alpha), the
    damped shifted-force approximation of the full Ewald sum)."""
    R, types, mol, L = water_box(nmol, seed)
    md = lib.system(1, 2, rc, skin, len(types), types, MASS, mol)
    kc = (ctypes.c_double * 2)(KCOUL, KCOUL)
    for t, pair in ((1, lambda: lib.EmDee_pair_lj_cut(EPS_OO, SIG_OO)), (2, lambda: lib.EmDee_pair_none())):
        models = (ctypes.c_void_p * 2)(pair(), pair())
        lib.EmDee_set_pair_multimodel(md, t, t, models, kc)
    if short == "dsf":
        second = lib.EmDee_shifted_force(lib.EmDee_coul_damped(short_alpha or 0.2))
    else:
        second = lib.EmDee_coul_damped(alpha)
    lib.EmDee_set_coul_multimodel(md, (ctypes.c_void_p * 2)(lib.EmDee_coul_long(), second))
    lib.EmDee_set_kspace_model(md, lib.EmDee_kspace_ewald(accuracy))
    md.Options.RotationMode = mode
    lib.upload(md, "charges", CHARGE[types - 1])
    lib.upload(md, "coordinates", R)
    lib.upload(md, "box", np.array([L]))
    return md, dict(R=R, types=types, mol=mol, L=L, N=len(types))


def kspace_labels(lib, md, N):
    """Per-atom k-space energy: atom energies of layer 1 minus layer 2."""
    lib.switch_model_layer(md, 1)
    e1 = lib.atom_energies(md, N)
    lib.switch_model_layer(md, 2)
    e2 = lib.atom_energies(md, N)
    lib.switch_model_layer(md, 1)
    return e1 - e2


class Surrogate:
    def __init__(self, lib, ntypes=2, rcut=6.0, seed=1, kcut=0.0):
        """kcut > 0 adds the exact long-range channel: the reciprocal Ewald
        sum over |k| <= kcut, the network learning the rest."""
        self.lib = lib
        self.h = lib.lib.EmDeeSurrogate_create(ntypes, rcut, seed)
        if kcut > 0.0:
            lib.lib.EmDeeSurrogate_set_long_range(self.h, kcut)

    def long_range_vectors(self):
        return self.lib.lib.EmDeeSurrogate_long_range_vectors(self.h)

    def sample(self, md, labels):
        labels = np.ascontiguousarray(labels, dtype=np.float64)
        self.lib.lib.EmDeeGPU_surrogate_sample(md, self.h, labels.ctypes.data_as(_dp))

    def rows(self):
        return self.lib.lib.EmDeeSurrogate_rows(self.h)

    def train(self, epochs, batch_rows=65536, lr=1e-3):
        return self.lib.lib.EmDeeSurrogate_train(self.h, epochs, batch_rows, lr)

    def params(self, ntypes=2):
        P = self.lib.lib.EmDeeSurrogate_param_count(self.h)
        p = np.zeros(P * ntypes, dtype=np.float32)
        din = 32 * ntypes
        norm = np.zeros(2 * din + 2 * ntypes, dtype=np.float32)
        self.lib.lib.EmDeeSurrogate_get_params(self.h, p.ctypes.data_as(_fp), norm.ctypes.data_as(_fp))
        return p, norm

    def set_params(self, p, norm):
        p = np.ascontiguousarray(p, dtype=np.float32)
        norm = np.ascontiguousarray(norm, dtype=np.float32)
        self.lib.lib.EmDeeSurrogate_set_params(self.h, p.ctypes.data_as(_fp), norm.ctypes.data_as(_fp))

    def attach(self, md, layer):
        self.lib.lib.EmDeeGPU_attach_surrogate(md, self.h, layer)

    def detach(self, md):
        self.lib.lib.EmDeeGPU_attach_surrogate(md, None, 0)

    def close(self):
        if self.h:
            self.lib.lib.EmDeeSurrogate_destroy(self.h)
            self.h = None


def collect(lib, md, surr, N, nsamples, stride, dt, kT=KB * 298.0, seed=11):
    """Exact-model dynamics (layer 1, velocity Verlet with a momentum refresh
    before every sample) and one labelled sample every `stride` steps."""
    lib.random_momenta(md, kT, True, seed)
    for k in range(nsamples):
        lib.switch_model_layer(md, 1)
        for _ in range(stride):
            lib.boost(md, 1.0, 0.0, 0.5 * dt)
            lib.displace(md, 1.0, 0.0, dt)
            lib.boost(md, 1.0, 0.0, 0.5 * dt)
        labels = kspace_labels(lib, md, N)
        surr.sample(md, labels)


def trajectory(lib, md, layer, nsteps, dt):
    lib.switch_model_layer(md, layer)
    for _ in range(nsteps):
        lib.boost(md, 1.0, 0.0, 0.5 * dt)
        lib.displace(md, 1.0, 0.0, dt)
        lib.boost(md, 1.0, 0.0, 0.5 * dt)


def exact_energy(lib, md):
    lib.switch_model_layer(md, 1)
    lib.compute_forces(md)
    return md.Energy.Potential


def propose(lib, md, proposal, nsteps, dt, respa_n=4):
    """One HMC trajectory of nsteps steps of dt.  proposal "exact": velocity
    Verlet on layer 1; "model": velocity Verlet on layer 2 (real-space forces
    plus whatever model is attached to it); "respa": r-RESPA with outer step dt
    on layer 1 and respa_n inner steps on layer 2 (mts.respa_step).  Returns
    the exact (layer-1) potential energy at the end."""
    from . import mts
    if proposal == "respa":
        for _ in range(nsteps):
            mts.respa_step(lib, md, dt, respa_n)
        return md.Energy.Potential          # the last boost evaluated layer 1 here
    layer = 1 if proposal == "exact" else 2
    trajectory(lib, md, layer, nsteps, dt)
    return md.Energy.Potential if layer == 1 else exact_energy(lib, md)


def hmc(lib, md, N, proposal, cycles, nsteps, dt, kT=KB * 298.0, seed=5, respa_n=4):
    """Hybrid Monte Carlo of the rigid water.  Every cycle draws canonical
    momenta (EmDee_random_momenta without the adjustment to a fixed kinetic
    energy), integrates one trajectory (`propose`: time-reversible and
    volume-preserving for every proposal kind) and accepts it with probability
    min(1, exp(-dH/kT)) on the exact layer-1 Hamiltonian, restoring the
    coordinates otherwise.  Returns acceptances, dH/kT per cycle, the exact
    potential-energy series and the wall time."""
    rng = np.random.default_rng(seed)
    md.Options.Compute = True
    U0 = exact_energy(lib, md)
    us, dhs, acc = [], [], 0
    lib.EmDeeGPU_synchronize(md)
    t0 = time.perf_counter()
    for c in range(cycles):
        R0 = lib.download(md, "coordinates", (N, 3))
        lib.random_momenta(md, kT, False, seed)
        K0 = md.Kinetic.Total
        U1 = propose(lib, md, proposal, nsteps, dt, respa_n)
        K1 = md.Kinetic.Total
        dH = (U1 + K1) - (U0 + K0)
        dhs.append(dH / kT)
        if dH <= 0.0 or rng.uniform() < math.exp(-dH / kT):
            acc += 1
            U0 = U1
        else:
            lib.upload(md, "coordinates", R0)
        us.append(U0)
    lib.EmDeeGPU_synchronize(md)
    return dict(accepted=acc, dH=np.array(dhs), U=np.array(us), seconds=time.perf_counter() - t0)


def eval_ms(lib, md, layer, n=20):
    """Wall milliseconds of one force evaluation of `layer` (GPU synchronised)."""
    lib.switch_model_layer(md, layer)
    lib.compute_forces(md)
    lib.EmDeeGPU_synchronize(md)
    t0 = time.perf_counter()
    for _ in range(n):
        lib.compute_forces(md)
    lib.EmDeeGPU_synchronize(md)
    return 1e3 * (time.perf_counter() - t0) / n


def zero_network(lib, kcut, ntypes=2, rcut=6.0):
    """A model whose network outputs zero: only the exact long-range channel."""
    s = Surrogate(lib, ntypes, rcut, seed=1, kcut=kcut)
    P = lib.lib.EmDeeSurrogate_param_count(s.h)
    din = 32 * ntypes
    s.set_params(np.zeros(P * ntypes, np.float32),
                 np.concatenate([np.zeros(din), np.ones(din), np.zeros(ntypes), np.ones(ntypes)]).astype(np.float32))
    return s


def force_fit(lib, md, N, mol, nmol, models, frames, stride, dt=1.0):
    """Held-out R^2 of the k-space forces (layer 1 minus layer 2) reproduced
    by each model on exact-dynamics frames it was not trained on: per atom and
    per molecule (the net force a rigid molecule feels)."""
    def molsum(F):
        B = np.zeros((nmol, 3))
        np.add.at(B, mol - 1, F)
        return B
    out = {k: [] for k in models}
    for _ in range(frames):
        trajectory(lib, md, 1, stride, dt)
        lib.switch_model_layer(md, 1)
        F1 = lib.download(md, "forces", (N, 3))
        lib.switch_model_layer(md, 2)
        F2 = lib.download(md, "forces", (N, 3))
        Fk = F1 - F2
        for name, m in models.items():
            m.attach(md, 2)
            lib.switch_model_layer(md, 2)
            dF = lib.download(md, "forces", (N, 3)) - F2
            m.detach(md)
            out[name].append((1.0 - np.var(Fk - dF) / np.var(Fk),
                              1.0 - np.var(molsum(Fk) - molsum(dF)) / np.var(molsum(Fk))))
    return {k: {"atom_force_r2": float(np.mean([v[0] for v in vs])),
                "molecule_force_r2": float(np.mean([v[1] for v in vs]))} for k, vs in out.items()}


def respa_error(lib, md, outer, nouter=4, dt=4.0, kT=KB * 298.0, seed=21):
    """r-RESPA 1:nouter (outer dt, inner dt/nouter on layer 2 plus whatever
    model is attached to it) from fresh momenta: RMS deviation of the exact
    total energy (layer-1 potential + kinetic) from its start, in kT."""
    from . import mts
    lib.random_momenta(md, kT, False, seed)
    lib.switch_model_layer(md, 1)
    lib.compute_forces(md)
    H0 = md.Energy.Potential + md.Kinetic.Total
    H = np.array(mts.respa(lib, md, outer, dt, nouter))
    return float(np.sqrt(np.mean((H - H0) ** 2)) / kT)


def water_hmc_workload(lib, nmol=4096, kcut=1.2, samples=60, stride=5, epochs=30, cycles=120, nsteps=10, dt=1.0,
                       seed=1, kT=KB * 298.0, log=None):
    """The learned-electrostatics workload on rigid SPC/E water: two layers
    (1: LJ + coul_long + Ewald, exact; 2: its real-space part), a model of the
    k-space part (exact reciprocal sum over |k| <= kcut plus a network trained
    on `samples` labelled configurations of plain-engine dynamics), used (a)
    as the proposal force of HMC with exact layer-1 Metropolis and (b) as the
    fast force of r-RESPA 1:4.  Compared with exact-force HMC and with the
    long-range channel alone (no network).  Returns a dict of measurements."""
    from .hmc import integrated_autocorrelation
    say = log or (lambda *a: None)
    md, cfg = two_layer_water(lib, nmol, seed)
    N, mol = cfg["N"], cfg["mol"]
    # equilibration from the lattice: 400 fs of exact dynamics, momenta
    # redrawn every 50 fs
    lib.random_momenta(md, kT, True, seed + 2)
    for k in range(8):
        trajectory(lib, md, 1, 50, dt)
        lib.random_momenta(md, kT, True, seed + 2)
    out = {"molecules": nmol, "atoms": N, "kcut": kcut, "dt_fs": dt, "nsteps": nsteps}
    out["exact_eval_ms"] = eval_ms(lib, md, 1)
    out["short_eval_ms"] = eval_ms(lib, md, 2)
    say(f"evaluations: exact {out['exact_eval_ms']:.3f} ms, real-space layer {out['short_eval_ms']:.3f} ms")
    net = Surrogate(lib, 2, 6.0, seed=seed, kcut=kcut)
    t0 = time.perf_counter()
    collect(lib, md, net, N, samples, stride, dt, kT, seed=seed + 10)
    t1 = time.perf_counter()
    loss = net.train(epochs, batch_rows=8 * N, lr=2e-3)
    lib.EmDeeGPU_synchronize(md)
    out.update(samples=samples, sample_seconds=t1 - t0, train_seconds=time.perf_counter() - t1, train_loss=loss,
               long_range_vectors=net.long_range_vectors())
    lowk = zero_network(lib, kcut)
    for name, m in (("model", net), ("long_range_only", lowk)):
        m.attach(md, 2)
        out[name + "_eval_ms"] = eval_ms(lib, md, 2)
        m.detach(md)
    say(f"model eval {out['model_eval_ms']:.3f} ms ({out['long_range_vectors']} k-vectors), trained {loss:.3f}")
    out["held_out"] = force_fit(lib, md, N, mol, nmol, {"model": net, "long_range_only": lowk}, 3, 20, dt)
    say(f"held-out k-space force R^2: {out['held_out']}")
    hmc_out = {}
    for name, m, kind in (("exact", None, "exact"), ("model", net, "model"), ("long_range_only", lowk, "model"),
                          ("real_space_only", None, "model")):
        if m is not None:
            m.attach(md, 2)
        n = cycles if name != "real_space_only" else max(4, cycles // 10)
        r = hmc(lib, md, N, kind, n, nsteps, dt, kT, seed=seed + 30)
        if m is not None:
            m.detach(md)
        tau = integrated_autocorrelation(r["U"]) if r["accepted"] > 1 else float(n)
        ess = n / max(tau, 1.0)
        x = np.exp(-np.clip(r["dH"], -50.0, 50.0))
        hmc_out[name] = {"cycles": n, "acceptance": r["accepted"] / n, "median_abs_dH_kT": float(np.median(np.abs(r["dH"]))),
                         "mean_exp_minus_dH": float(x.mean()), "tau_cycles": float(tau),
                         "ms_per_cycle": 1e3 * r["seconds"] / n, "effective_samples_per_s": ess / r["seconds"]}
        say(f"hmc {name}: {hmc_out[name]}")
    hmc_out["model_over_exact_ess"] = hmc_out["model"]["effective_samples_per_s"] / hmc_out["exact"]["effective_samples_per_s"]
    out["hmc"] = hmc_out
    net.attach(md, 2)
    out["respa_model_fast_energy_rms_kT"] = respa_error(lib, md, 25)
    net.detach(md)
    out["respa_real_space_fast_energy_rms_kT"] = respa_error(lib, md, 25)
    say(f"r-RESPA 1:4 (4 fs) energy RMS: model fast {out['respa_model_fast_energy_rms_kT']:.3f} kT, "
        f"real-space fast {out['respa_real_space_fast_energy_rms_kT']:.3f} kT")
    net.close()
    lowk.close()
    lib.EmDeeGPU_destroy(ctypes.byref(md))
    return out


def evaluation_costs(lib, nmol=4096, kcut=1.2, evals=20, seed=1):
    """Force-evaluation costs on the two-layer water box, for kernel traces:
    `evals` evaluations each of the exact layer 1, of the real-space layer 2,
    and of layer 2 with a model attached (long-range channel plus network; the
    network's weights are zero, which changes no kernel's work).  Returns the
    wall milliseconds per evaluation of each."""
    md, cfg = two_layer_water(lib, nmol, seed)
    out = {"molecules": nmol, "atoms": cfg["N"], "kcut": kcut,
           "exact_eval_ms": eval_ms(lib, md, 1, evals), "short_eval_ms": eval_ms(lib, md, 2, evals)}
    m = zero_network(lib, kcut)
    m.attach(md, 2)
    out["model_eval_ms"] = eval_ms(lib, md, 2, evals)
    out["long_range_vectors"] = m.long_range_vectors()
    m.detach(md)
    m.close()
    lib.EmDeeGPU_destroy(ctypes.byref(md))
    return out


if __name__ == "__main__":
    # Workload: python -m emdee_amd.surrogate [molecules] [kcut]
    # Evaluation costs only: python -m emdee_amd.surrogate --evals [molecules] [kcut] [evaluations]
    import json
    import os
    import sys

    from . import load
    args = sys.argv[1:]
    costs = bool(args) and args[0] == "--evals"
    if costs:
        args = args[1:]
    nmol = int(args[0]) if len(args) > 0 else 4096
    kc = float(args[1]) if len(args) > 1 else 1.2
    os.environ.setdefault("EMDEE_QUIET", "1")
    if costs:
        res = evaluation_costs(load(), nmol, kc, int(args[2]) if len(args) > 2 else 20)
    else:
        res = water_hmc_workload(load(), nmol, kc, log=lambda *a: print(*a, flush=True))
    print(json.dumps(res), flush=True)
