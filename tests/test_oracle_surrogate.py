"""Pins of the FP64 restatement of the learned energy model
(tests/surrogate_ref.py), which checks surrogate.hip: analytic forces equal
-grad E by central differences, descriptors are translation invariant and
vanish smoothly at the cutoff."""
import numpy as np
import pytest

import surrogate_ref as sr
from emdee_amd.water_box import water_box


def small_water():
    R, types, mol, L = water_box(27, seed=4)
    return R, types - 1, mol, L


def norm_for(D, nt):
    din = nt * sr.NB
    return np.concatenate([D.mean(0), D.std(0) + 1e-3, np.array([0.1, -0.2]), np.array([2.0, 0.5])])


def test_forces_are_minus_gradient():
    R, types, mol, L = small_water()
    rc = 4.0
    D = sr.descriptors(R, types, L, rc, mol, 2)
    norm = norm_for(D, 2)
    p = sr.random_params(2, seed=1)
    E, F, _, _ = sr.energy_forces(p, norm, R, types, L, rc, mol, 2)
    h = 1e-5
    rng = np.random.default_rng(0)
    for a in rng.choice(len(R), 6, replace=False):
        for x in range(3):
            Rp, Rm = R.copy(), R.copy()
            Rp[a, x] += h
            Rm[a, x] -= h
            Ep = sr.energy_forces(p, norm, Rp, types, L, rc, mol, 2)[0]
            Em = sr.energy_forces(p, norm, Rm, types, L, rc, mol, 2)[0]
            assert F[a, x] == pytest.approx(-(Ep - Em) / (2 * h), rel=1e-5, abs=1e-8)
    assert np.allclose(F.sum(0), 0.0, atol=1e-9)


def test_descriptor_invariances():
    R, types, mol, L = small_water()
    rc = 4.0
    D = sr.descriptors(R, types, L, rc, mol, 2)
    shift = np.array([1.3, -2.1, 0.7])
    assert np.allclose(sr.descriptors(R + shift, types, L, rc, mol, 2), D, atol=1e-12)
    # a pair at the cutoff contributes nothing
    mu, eta = sr.constants(rc)
    assert np.all(0.5 * (np.cos(np.pi * rc / rc) + 1.0) * np.exp(-eta * (rc - mu) ** 2) < 1e-15)
