"""Pins of the counter-based generator of the batched HMC driver and of the
reference-API HMC replay it feeds:

* splitmix64 (Steele, Lea & Flood 2014) against its published first outputs
  for state 0;
* uniforms in (0, 1) and Gaussian momenta with the right moments;
* the product library's host copy of the generator equals the oracle's
  restatement (the device uses the same formula);
* HMC with an exact-energy Metropolis test satisfies <exp(-dH/kT)> = 1
  (a property of volume-preserving, time-reversible integrators).
"""
import ctypes
import math

import numpy as np
import pytest

import hmc_ref
from emdee_amd.hmc import replica_set


def test_splitmix64_known_answers(oracle):
    f = oracle.lib.oracle_splitmix64_next
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
    st = ctypes.c_uint64(0)
    assert f(ctypes.byref(st)) == 0xE220A8397B1DCDAF
    assert f(ctypes.byref(st)) == 0x6E789E6AA1B965F4


def test_momenta_statistics(oracle):
    P = hmc_ref.momenta(oracle, 99, 3, 7, 50000, 1.7)
    assert abs(P.mean()) < 5 * math.sqrt(1.7 / P.size)
    assert P.var() == pytest.approx(1.7, rel=0.02)
    assert np.mean((P / math.sqrt(1.7)) ** 4) == pytest.approx(3.0, rel=0.05)
    u = np.array([hmc_ref.accept_uniform(oracle, 99, r, c) for r in range(40) for c in range(50)])
    assert 0.0 < u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 5 * math.sqrt(1 / 12 / u.size)


def test_product_host_generator_matches_oracle(gpu, oracle):
    """Host code of the product library only: no device call."""
    lib = gpu.lib
    P = np.zeros((100, 3))
    lib.EmDeeHMC_momenta(2024, 5, 11, 100, 0.9, P.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    Q = hmc_ref.momenta(oracle, 2024, 5, 11, 100, 0.9)
    assert np.max(np.abs(P - Q)) <= 1e-15 * np.max(np.abs(Q))
    assert lib.EmDeeHMC_uniform(2024, 5, 11) == hmc_ref.accept_uniform(oracle, 2024, 5, 11)


def test_hmc_identity(oracle):
    """After equilibration, <exp(-dH/kT)> = 1 over the trajectories."""
    N, L, T, Rs = replica_set(4, seed=5, nmin=108, nmax=160)
    _, _, acc, dH = hmc_ref.replay(oracle, oracle, N, L, T, Rs, 60, dt=0.005, nsteps=10, seed=11)
    w = np.exp(-dH[:, 30:] / T[:, None])
    assert np.mean(w) == pytest.approx(1.0, abs=0.05)
    assert 0.5 < acc.sum() / acc.size / 60 <= 1.0
