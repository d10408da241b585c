"""Pins of the oracle's EmDee_verlet_step (EmDeeCode.f90:1107-1211):

* its trajectory is the boost(dt/2) / move(dt) / boost(dt/2) sequence of the
  separate entry points, bit for bit;
* the shadow Hamiltonian it reports (ShadowPotential + ShadowKinetic) is the
  modified energy that velocity Verlet conserves to higher order: its
  fluctuation along a trajectory is far below that of the plain energy, for
  free atoms and for rigid bodies, and it tends to the plain energy as dt->0.
"""
import numpy as np
import pytest

import hosts
import water


def trajectory(lib, md, nsteps, dt):
    H, Hs = [], []
    for _ in range(nsteps):
        lib.verlet_step(md, dt)
        H.append(md.Energy.Potential + md.Kinetic.Total)
        Hs.append(md.Energy.ShadowPotential + md.Kinetic.ShadowKinetic)
    return np.array(H), np.array(Hs)


def test_matches_separate_steps(oracle):
    a, cfg = hosts.lj_sample_system(oracle, "lj_sf")
    b, _ = hosts.lj_sample_system(oracle, "lj_sf")
    for _ in range(30):
        oracle.verlet_step(a, 0.005)
        oracle.boost(b, 1.0, 0.0, 0.0025)
        oracle.displace(b, 1.0, 0.0, 0.005)
        oracle.boost(b, 1.0, 0.0, 0.0025)
    N = cfg["N"]
    assert np.array_equal(oracle.download(a, "coordinates", (N, 3)), oracle.download(b, "coordinates", (N, 3)))
    assert np.array_equal(oracle.download(a, "momenta", (N, 3)), oracle.download(b, "momenta", (N, 3)))
    assert a.Energy.Potential == b.Energy.Potential


def test_shadow_hamiltonian_free_atoms(oracle):
    """For velocity Verlet, H - H_shadow = O(dt^2) and the shadow energy
    fluctuates as O(dt^3) along a trajectory (vs O(dt^2) for H): halving dt
    divides the first by ~4 and the fluctuation of the second by ~8."""
    stats = []
    for dt in (0.005, 0.0025):
        md, cfg = hosts.lj_sample_system(oracle, "lj_sf")
        H, Hs = trajectory(oracle, md, int(round(3.0 / dt)), dt)
        stats.append((np.mean(Hs - H), np.std(Hs), np.std(H)))
    (d1, s1, h1), (d2, s2, h2) = stats
    assert 3.5 < d1 / d2 < 4.5
    assert s1 / s2 > 6.0
    assert s2 < 0.1 * h2


def test_shadow_hamiltonian_rigid_bodies(oracle):
    """H - H_shadow = O(dt^2) with rigid-body rotation terms included."""
    d = []
    for dt in (1.0, 0.5):
        md, cfg = water.spce_fixture_system(oracle, rc=9.0, skin=1.0)
        oracle.random_momenta(md, water.KB * 298.0, True, 86245)
        H, Hs = trajectory(oracle, md, int(round(30.0 / dt)), dt)
        d.append(np.mean(Hs - H))
        assert md.Kinetic.ShadowRotational != 0.0
    assert 3.5 < d[0] / d[1] < 4.5


def test_rdf_against_histogram(oracle):
    """EmDee_rdf of an LJ fluid = normalised histogram of all pair distances
    below the cutoff (no exclusions: the list holds every pair within rc+skin)."""
    import ctypes
    md, cfg = hosts.lj_sample_system(oracle, "lj_cut")
    bins, rc = 50, 2.5
    g = np.zeros(bins)
    one = np.array([1], np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    oracle.EmDee_rdf(md, bins, rc, 1, one.ctypes.data_as(ip), one.ctypes.data_as(ip),
                     g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    R = oracle.download(md, "coordinates", (cfg["N"], 3))
    L, N = cfg["L"], cfg["N"]
    d = R[:, None, :] - R[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(-1))[np.triu_indices(N, 1)]
    counts, _ = np.histogram(r[r < rc], bins=bins, range=(0, rc))
    shell = 4 / 3 * np.pi * ((np.arange(1, bins + 1) * rc / bins) ** 3 - (np.arange(bins) * rc / bins) ** 3)
    expected = 2 * counts / shell / (N * N / L ** 3)
    # at most one pair may change bin by rounding at a bin edge
    one_count = 2.0 / (shell * N * N / L ** 3)
    assert np.all(np.abs(g - expected) <= one_count + 1e-9 * expected)
    assert np.sum(np.abs(g - expected) > 1e-9 * expected) <= 2
    assert g[-10:].mean() == pytest.approx(1.0, abs=0.1)
