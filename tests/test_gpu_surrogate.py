"""The learned energy model on the GPU (surrogate.hip) against its FP64
restatement (tests/surrogate_ref.py): descriptors, the BF16-MFMA network and
its backward pass, forces, the exact long-range channel (reciprocal sum over
|k| <= kcut), the first Adam step of training and the fit of held-out labels.
What the trained model gains as a proposal force is in test_gpu_water_hmc.py."""
import numpy as np
import pytest

import surrogate_ref as sr
from emdee_amd import surrogate as sg
from emdee_amd.water_box import KB, MASS

pytestmark = pytest.mark.gpu

NMOL, RC = 512, 6.0


def layer2_state(gpu, md, N):
    gpu.switch_model_layer(md, 2)
    gpu.compute_forces(md)
    return gpu.download(md, "forces", (N, 3)), md.Energy.Coulomb


def lowk_reference(cfg, R, kcut):
    q = sg.CHARGE[cfg["types"] - 1]
    return sr.lowk_energy_forces(R, q, cfg["L"], 0.3, kcut, sg.KCOUL)


@pytest.mark.parametrize("kcut", [0.0, 0.9])
def test_forward_and_forces_match_fp64(gpu, kcut):
    md, cfg = sg.two_layer_water(gpu, NMOL, seed=3)
    N = cfg["N"]
    F0, E0 = layer2_state(gpu, md, N)
    R = gpu.download(md, "coordinates", (N, 3))
    types, mol, L = cfg["types"] - 1, cfg["mol"], cfg["L"]
    D = sr.descriptors(R, types, L, RC, mol, 2)
    norm = np.concatenate([D.mean(0), D.std(0) + 1e-3, [0.02, -0.01], [0.05, 0.03]]).astype(np.float32)
    p = sr.random_params(2, seed=7)
    s = sg.Surrogate(gpu, 2, RC, seed=1, kcut=kcut)
    s.set_params(p, norm)
    s.attach(md, 2)
    F1, E1 = layer2_state(gpu, md, N)
    e_atoms = gpu.atom_energies(md, N)
    s.detach(md)
    e_base = gpu.atom_energies(md, N)
    E_ref, F_ref, e_ref, _ = sr.energy_forces(p.astype(np.float64), norm.astype(np.float64), R, types, L, RC, mol, 2)
    if kcut > 0.0:
        El, Fl, el, _ = lowk_reference(cfg, R, kcut)
        E_ref, F_ref, e_ref = E_ref + El, F_ref + Fl, e_ref + el
    dF = F1 - F0
    ferr = np.linalg.norm(dF - F_ref) / np.linalg.norm(F_ref)
    print("surrogate: energy", E1 - E0, E_ref, "force rel rms", ferr, flush=True)
    assert E1 - E0 == pytest.approx(E_ref, rel=2e-2, abs=1e-3 * np.abs(e_ref).sum())
    assert ferr < 3e-2
    assert np.max(np.abs((e_atoms - e_base) - e_ref)) < 3e-2 * np.max(np.abs(e_ref))
    s.close()


def test_long_range_channel_is_the_partial_ewald_sum(gpu):
    """With a network that outputs zero, the attached model is exactly the
    reciprocal sum over |k| <= kcut: energy, forces, atom energies and virial
    against the FP64 restatement (pinned to the oracle's Ewald sum)."""
    kcut = 1.0
    md, cfg = sg.two_layer_water(gpu, NMOL, seed=6)
    N = cfg["N"]
    F0, _ = layer2_state(gpu, md, N)
    E0, W0 = md.Energy.Potential, md.Virial
    e0 = gpu.atom_energies(md, N)
    R = gpu.download(md, "coordinates", (N, 3))
    s = sg.Surrogate(gpu, 2, RC, seed=1, kcut=kcut)
    P = len(sr.random_params(2, seed=1)) // 2
    din = 2 * sr.NB
    s.set_params(np.zeros(2 * P, np.float32),
                 np.concatenate([np.zeros(din), np.ones(din), [0.0, 0.0], [1.0, 1.0]]).astype(np.float32))
    s.attach(md, 2)
    F1, _ = layer2_state(gpu, md, N)
    E1, W1 = md.Energy.Potential, md.Virial
    e1 = gpu.atom_energies(md, N)
    El, Fl, el, Wl = lowk_reference(cfg, R, kcut)
    print("long-range channel: vectors", s.long_range_vectors(), "E", E1 - E0, El, "W", W1 - W0, Wl, flush=True)
    assert s.long_range_vectors() == len(sr.lowk_vectors(cfg["L"], kcut))
    assert E1 - E0 == pytest.approx(El, rel=1e-9, abs=1e-12 * abs(E0))
    assert np.max(np.abs((F1 - F0) - Fl)) <= 1e-9 * np.max(np.abs(Fl))
    assert np.max(np.abs((e1 - e0) - el)) <= 1e-9 * np.max(np.abs(el))
    # the added forces also change the rigid-body virial, -sum_k F_k . delta_k
    # (EmDeeData.f90:926-953), delta_k the atom's offset from its molecule's centre of mass
    m = sg.MASS[cfg["types"] - 1]
    delta = np.zeros_like(R)
    for b in range(NMOL):
        idx = np.arange(3 * b, 3 * b + 3)
        X = R[idx].copy()
        X[1:] -= cfg["L"] * np.round((X[1:] - X[0]) / cfg["L"])
        delta[idx] = X - (m[idx, None] * X).sum(0) / m[idx].sum()
    assert W1 - W0 == pytest.approx(Wl - np.sum(Fl * delta), rel=1e-8, abs=1e-12 * abs(W0))
    s.close()


def test_descriptor_cutoff_above_the_system_cutoff_is_rejected(gpu):
    """The compact list follows the engine's Verlet list, which is complete
    only to the system's cutoff: a larger descriptor cutoff is an error."""
    import subprocess
    import sys
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = (
        "import sys; sys.path[:0] = [{root!r}]\n"
        "import emdee_amd\n"
        "from emdee_amd import surrogate as sg\n"
        "lib = emdee_amd.load()\n"
        "md, cfg = sg.two_layer_water(lib, 216, seed=1, rc=7.0, skin=1.0)\n"
        "s = sg.Surrogate(lib, 2, 8.0, seed=1)\n"
        "s.attach(md, 2)\n"
        "print('not reached')\n"
    ).format(root=root)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, EMDEE_QUIET="1"))
    assert r.returncode == 1, r.stderr
    assert "Error in surrogate attachment: the descriptor cutoff exceeds the system's cutoff." in r.stderr


def test_first_adam_step_follows_gradient(gpu):
    md, cfg = sg.two_layer_water(gpu, NMOL, seed=4)
    N = cfg["N"]
    s = sg.Surrogate(gpu, 2, RC, seed=2)
    Ds, ys = [], []
    types, mol, L = cfg["types"] - 1, cfg["mol"], cfg["L"]
    gpu.random_momenta(md, KB * 298.0, True, 3)
    for k in range(3):
        sg.trajectory(gpu, md, 1, 3, 1.0)
        y = sg.kspace_labels(gpu, md, N)
        s.sample(md, y)
        R = gpu.download(md, "coordinates", (N, 3))
        Ds.append(sr.descriptors(R, types, L, RC, mol, 2))
        ys.append(y)
    p0, _ = s.params()
    s.train(1, batch_rows=3 * N, lr=1e-5)
    p1, norm = s.params()
    D, y, t = np.concatenate(Ds), np.concatenate(ys), np.tile(types, 3)
    g = sr.loss_grad(p0.astype(np.float64), norm.astype(np.float64), D, t, y, 2)
    big = np.abs(g) > 1e-2 * np.max(np.abs(g))
    agree = np.mean(np.sign(p0 - p1)[big] == np.sign(g)[big])
    print("sign agreement", agree, "of", big.sum())
    assert agree > 0.95
    s.close()


def test_training_fits_held_out_labels(gpu):
    """Trained on k-space atom energies of 40 configurations, the model
    explains a clear part of their variance on later configurations it has
    not seen (the untrained model explains none of it)."""
    md, cfg = sg.two_layer_water(gpu, NMOL, seed=5)
    N = cfg["N"]
    s = sg.Surrogate(gpu, 2, RC, seed=3)
    sg.collect(gpu, md, s, N, nsamples=40, stride=4, dt=1.0)
    loss = s.train(40, batch_rows=8 * N, lr=2e-3)
    types = cfg["types"] - 1
    r2 = []
    for k in range(3):
        sg.trajectory(gpu, md, 1, 8, 1.0)
        y = sg.kspace_labels(gpu, md, N)
        gpu.switch_model_layer(md, 2)
        e0 = gpu.atom_energies(md, N)
        s.attach(md, 2)
        pred = gpu.atom_energies(md, N) - e0
        s.detach(md)
        for t in (0, 1):
            sel = types == t
            r2.append(1.0 - np.var(pred[sel] - y[sel]) / np.var(y[sel]))
    print("training loss", loss, "held-out R^2 per species/config", np.round(r2, 3))
    assert loss < 0.8
    assert np.mean(r2) > 0.2
    s.close()
