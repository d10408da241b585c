"""Rigid bodies on the GPU against the oracle (ArBee.f90 + the body parts of
EmDeeData.f90), on the NIST SPC/E fixture set up as the reference's rigid
body tests do (test/test_rigid_body_exact.f90, _miller.f90): body set-up and
principal axes, forces/energies/virials including the body virial, random
momenta, exact and NO_SQUISH rotation dynamics, and the body views of
EmDee_download."""
import numpy as np
import pytest

import water

pytestmark = pytest.mark.gpu


def pair(gpu, oracle, **kw):
    return water.spce_fixture_system(gpu, **kw), water.spce_fixture_system(oracle, **kw)


def test_setup_and_forces(gpu, oracle):
    (mg, cfg), (mo, _) = pair(gpu, oracle)
    N, nb = cfg["N"], cfg["N"] // 3
    for view, shape, tol in [("inertia", (nb, 3), 1e-12), ("centersOfMass", (nb, 3), 1e-13),
                             ("coordinates", (N, 3), 1e-13)]:
        a, b = gpu.download(mg, view, shape), oracle.download(mo, view, shape)
        assert np.max(np.abs(a - b)) <= tol * np.max(np.abs(b)), view
    qg, qo = gpu.download(mg, "quaternions", (nb, 4)), oracle.download(mo, "quaternions", (nb, 4))
    # principal axes come from an eigensolver: a rounding difference in the
    # inertia tensor moves an eigenvector by about eps/(relative eigengap), so
    # nearly symmetric bodies set the bound, not the typical 1e-15 agreement
    assert np.max(np.abs(qg - qo)) < 1e-10
    assert np.median(np.abs(qg - qo)) < 1e-14
    Fg, Fo = gpu.download(mg, "forces", (N, 3)), oracle.download(mo, "forces", (N, 3))
    assert np.max(np.abs(Fg - Fo)) <= 1e-12 * np.max(np.abs(Fo))
    assert mg.Energy.Potential == pytest.approx(mo.Energy.Potential, rel=1e-12)
    assert mg.Energy.Coulomb == pytest.approx(mo.Energy.Coulomb, rel=1e-12)
    assert mg.BodyVirial == pytest.approx(mo.BodyVirial, rel=1e-11)
    assert mg.Virial == pytest.approx(mo.Virial, rel=1e-11)
    assert (mg.DoF, mg.RotDoF) == (mo.DoF, mo.RotDoF) == (6 * nb - 3, 3 * nb)


def test_random_momenta_and_views(gpu, oracle):
    (mg, cfg), (mo, _) = pair(gpu, oracle)
    N, nb = cfg["N"], cfg["N"] // 3
    for lib, md in ((gpu, mg), (oracle, mo)):
        lib.random_momenta(md, water.KB * 298.0, True, 86245)
    assert mg.Kinetic.Total == pytest.approx(mo.Kinetic.Total, rel=1e-13)
    assert mg.Kinetic.Rotational == pytest.approx(mo.Kinetic.Rotational, rel=1e-12)
    for view, shape in [("momenta", (N, 3)), ("bodymom", (nb, 3)), ("angmom", (nb, 3)), ("quatmom", (nb, 4))]:
        a, b = gpu.download(mg, view, shape), oracle.download(mo, view, shape)
        assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b)), view
    # one kick fills forces and torques
    for lib, md in ((gpu, mg), (oracle, mo)):
        lib.boost(md, 1.0, 0.0, 1.0)
    for view, shape in [("bodyforces", (nb, 3)), ("torques", (nb, 3)), ("quattau", (nb, 4))]:
        a, b = gpu.download(mg, view, shape), oracle.download(mo, view, shape)
        assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b)), view
    assert mg.Kinetic.Total == pytest.approx(mo.Kinetic.Total, rel=1e-11)


@pytest.mark.parametrize("mode", [0, 3])
def test_dynamics(gpu, oracle, mode):
    (mg, cfg), (mo, _) = pair(gpu, oracle, mode=mode)
    N = cfg["N"]
    for lib, md in ((gpu, mg), (oracle, mo)):
        lib.random_momenta(md, water.KB * 298.0, True, 86245)
        water.run(lib, md, 20, 1.0)
    assert mg.Builds == mo.Builds
    H = lambda md: md.Energy.Potential + md.Kinetic.Total
    assert H(mg) == pytest.approx(H(mo), rel=1e-9)
    assert mg.Virial == pytest.approx(mo.Virial, rel=1e-7)
    Rg, Ro = gpu.download(mg, "coordinates", (N, 3)), oracle.download(mo, "coordinates", (N, 3))
    assert np.max(np.abs(Rg - Ro)) < 1e-9


def test_momenta_roundtrip(gpu, oracle):
    (mg, cfg), (mo, _) = pair(gpu, oracle)
    N = cfg["N"]
    P = np.random.default_rng(3).standard_normal((N, 3)) * 0.01
    for lib, md in ((gpu, mg), (oracle, mo)):
        lib.upload(md, "momenta", P)
    assert mg.Kinetic.Total == pytest.approx(mo.Kinetic.Total, rel=1e-12)
    # the download rebuilds member momenta from the body's omega and principal
    # axes, so it inherits the eigenvector bound of the quaternions
    a, b = gpu.download(mg, "momenta", (N, 3)), oracle.download(mo, "momenta", (N, 3))
    assert np.max(np.abs(a - b)) <= 1e-10 * np.max(np.abs(b))
