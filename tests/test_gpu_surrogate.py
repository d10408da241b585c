"""The learned energy model on the GPU (surrogate.hip) against its FP64
restatement (tests/surrogate_ref.py): descriptors, BF16-MFMA forward pass,
forces from the backward pass, the first Adam step of training, and what the
trained model gains as an HMC proposal force for rigid water."""
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


def test_forward_and_forces_match_fp64(gpu):
    md, cfg = sg.two_layer_water(gpu, NMOL, seed=3)
    N = cfg["N"]
    F0, E0 = layer2_state(gpu, md, N)
    R = gpu.download(md, "coordinates", (N, 3))
    types, mol, L = cfg["types"] - 1, cfg["mol"], cfg["L"]
    D = sr.descriptors(R, types, L, RC, mol, 2)
    norm = np.concatenate([D.mean(0), D.std(0) + 1e-3, [0.02, -0.01], [0.05, 0.03]]).astype(np.float32)
    p = sr.random_params(2, seed=7)
    s = sg.Surrogate(gpu, 2, RC, seed=1)
    s.set_params(p, norm)
    s.attach(md, 2)
    F1, E1 = layer2_state(gpu, md, N)
    e_atoms = gpu.atom_energies(md, N)
    s.detach(md)
    e_base = gpu.atom_energies(md, N)
    E_ref, F_ref, e_ref, _ = sr.energy_forces(p.astype(np.float64), norm.astype(np.float64), R, types, L, RC, mol, 2)
    dF = F1 - F0
    ferr = np.linalg.norm(dF - F_ref) / np.linalg.norm(F_ref)
    print("surrogate: energy", E1 - E0, E_ref, "force rel rms", ferr)
    assert E1 - E0 == pytest.approx(E_ref, rel=2e-2, abs=1e-3 * np.abs(e_ref).sum())
    assert ferr < 3e-2
    assert np.max(np.abs((e_atoms - e_base) - e_ref)) < 3e-2 * np.max(np.abs(e_ref))
    s.close()


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
