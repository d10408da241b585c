"""EmDee_verlet_step with its shadow energies (EmDeeCode.f90:1107-1211) and
EmDee_rdf (EmDeeCode.f90:1281-1395) on the GPU against the oracle."""
import ctypes

import numpy as np
import pytest

import hosts
import water

pytestmark = pytest.mark.gpu


def run(lib, md, n, dt):
    for _ in range(n):
        lib.verlet_step(md, dt)


def test_verlet_lj(gpu, oracle):
    mg, cfg = hosts.lj_sample_system(gpu, "lj_sf")
    mo, _ = hosts.lj_sample_system(oracle, "lj_sf")
    run(gpu, mg, 50, 0.005)
    run(oracle, mo, 50, 0.005)
    assert mg.Builds == mo.Builds
    for a, b in [(mg.Energy.ShadowPotential, mo.Energy.ShadowPotential),
                 (mg.Kinetic.ShadowKinetic, mo.Kinetic.ShadowKinetic),
                 (mg.Kinetic.Total, mo.Kinetic.Total), (mg.Energy.Potential, mo.Energy.Potential)]:
        assert a == pytest.approx(b, rel=1e-10)
    N = cfg["N"]
    assert np.max(np.abs(gpu.download(mg, "coordinates", (N, 3)) - oracle.download(mo, "coordinates", (N, 3)))) < 1e-10


@pytest.mark.parametrize("mode", [0, 3])
def test_verlet_rigid_water(gpu, oracle, mode):
    mg, cfg = water.spce_fixture_system(gpu, rc=9.0, skin=1.0, mode=mode)
    mo, _ = water.spce_fixture_system(oracle, rc=9.0, skin=1.0, mode=mode)
    for lib, md in ((gpu, mg), (oracle, mo)):
        lib.random_momenta(md, water.KB * 298.0, True, 86245)
        run(lib, md, 20, 1.0)
    for a, b in [(mg.Energy.ShadowPotential, mo.Energy.ShadowPotential),
                 (mg.Kinetic.ShadowKinetic, mo.Kinetic.ShadowKinetic),
                 (mg.Kinetic.ShadowRotational, mo.Kinetic.ShadowRotational),
                 (mg.Kinetic.Rotational, mo.Kinetic.Rotational)]:
        assert a == pytest.approx(b, rel=1e-8)


def rdf(lib, md, bins, rc, pairs):
    it = np.array([p[0] for p in pairs], np.int32)
    jt = np.array([p[1] for p in pairs], np.int32)
    g = np.zeros(bins * len(pairs))
    ip = ctypes.POINTER(ctypes.c_int)
    lib.EmDee_rdf(md, bins, rc, len(pairs), it.ctypes.data_as(ip), jt.ctypes.data_as(ip),
                  g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return g.reshape(len(pairs), bins)


def test_rdf_water(gpu, oracle):
    mg, cfg = water.spce_fixture_system(gpu, rc=9.0, skin=1.0)
    mo, _ = water.spce_fixture_system(oracle, rc=9.0, skin=1.0)
    pairs = [(1, 1), (1, 2), (2, 2)]
    assert np.array_equal(rdf(gpu, mg, 90, 9.0, pairs), rdf(oracle, mo, 90, 9.0, pairs))


def test_rdf_lj(gpu, oracle):
    mg, cfg = hosts.lj_sample_system(gpu, "lj_cut")
    mo, _ = hosts.lj_sample_system(oracle, "lj_cut")
    run(gpu, mg, 30, 0.005)
    run(oracle, mo, 30, 0.005)
    assert np.array_equal(rdf(gpu, mg, 100, 3.0, [(1, 1)]), rdf(oracle, mo, 100, 3.0, [(1, 1)]))
