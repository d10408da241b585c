"""Learned stand-in for the k-space electrostatics of rigid water
(EmDeeSurrogate_* and EmDeeGPU_attach_surrogate of include/emdee_gpu.h).

Two model layers describe the same water (reference mechanism for multiple
time steps, EmDeeCode.f90:273-305, 486-520, 929-946):
  layer 1: the exact model, LJ + coul_long + kspace_ewald;
  layer 2: LJ + coul_damped(alpha), i.e. coul_long's real-space part alone.
Their per-atom energy difference is the k-space (reciprocal + self) energy
of each atom; the surrogate learns it from descriptors of the atom's 6 A
environment and, attached to layer 2, restores an approximation of the
exact model at the cost of the real-space layer.  Exact layer-1 energies
judge every Metropolis acceptance in `hmc`.
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
    def __init__(self, lib, ntypes=2, rcut=6.0, seed=1):
        self.lib = lib
        self.h = lib.lib.EmDeeSurrogate_create(ntypes, rcut, seed)

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


def hmc(lib, md, N, proposal_layer, cycles, nsteps, dt, kT=KB * 298.0, seed=5):
    """Hybrid Monte Carlo of the rigid water: momenta from EmDee_random_momenta,
    an nsteps trajectory on `proposal_layer` (1: exact forces; 2: real-space
    forces, plus the surrogate when one is attached), Metropolis on the exact
    layer-1 energy.  Returns (accepted, U series, seconds)."""
    rng = np.random.default_rng(seed)
    U0 = exact_energy(lib, md)
    us, acc = [], 0
    t0 = time.perf_counter()
    for c in range(cycles):
        R0 = lib.download(md, "coordinates", (N, 3))
        lib.random_momenta(md, kT, True, seed + c)
        K0 = md.Kinetic.Total
        trajectory(lib, md, proposal_layer, nsteps, dt)
        K1 = md.Kinetic.Total
        U1 = exact_energy(lib, md)
        dH = (U1 + K1) - (U0 + K0)
        if dH <= 0.0 or rng.uniform() < math.exp(-dH / kT):
            acc += 1
            U0 = U1
        else:
            lib.upload(md, "coordinates", R0)
            U0 = exact_energy(lib, md)
        us.append(U0)
    return acc, np.array(us), time.perf_counter() - t0


if __name__ == "__main__":
    # Profiling workload: python -m emdee_amd.surrogate [molecules] [samples]
    import os
    import sys

    from . import load
    nmol = int(sys.argv[1]) if len(sys.argv) > 1 else 4096
    nsamples = int(sys.argv[2]) if len(sys.argv) > 2 else 20
    os.environ.setdefault("EMDEE_QUIET", "1")
    lib = load()
    md, cfg = two_layer_water(lib, nmol, seed=1)
    N = cfg["N"]
    s = Surrogate(lib, 2, 6.0, seed=1)
    t0 = time.perf_counter()
    collect(lib, md, s, N, nsamples=nsamples, stride=2, dt=1.0)
    t1 = time.perf_counter()
    loss = s.train(5, batch_rows=8 * N, lr=2e-3)
    t2 = time.perf_counter()
    s.attach(md, 2)
    trajectory(lib, md, 2, 10, 1.0)
    lib.EmDeeGPU_synchronize(md)
    t3 = time.perf_counter()
    print(f"{N} atoms: {nsamples} labelled samples {t1 - t0:.2f} s, 5 epochs {t2 - t1:.2f} s (loss {loss:.3f}), "
          f"10 surrogate steps {t3 - t2:.3f} s", flush=True)
    s.close()
