"""Ewald sums and bonded terms on the GPU against the oracle
(kspace_ewald.f90:188-314, modelClass_kspace.f90:265-334,
EmDeeData.f90:443-550, EmDeeCode.f90:1248-1252), on the configurations of
the reference's coul_long test (test/test_coul_long.f90: rigid SPC/E with
kspace_ewald(1e-4)), on rock salt, and on flexible water with harmonic
bonds and angles (with and without the k-space discounts)."""
import numpy as np
import pytest

import ionic
import water

pytestmark = pytest.mark.gpu


def compare(gpu, oracle, mg, mo, N, ftol=1e-11, etol=1e-12):
    Fg, Fo = gpu.download(mg, "forces", (N, 3)), oracle.download(mo, "forces", (N, 3))
    assert np.max(np.abs(Fg - Fo)) <= ftol * np.max(np.abs(Fo))
    for f in ("Potential", "Coulomb", "Dispersion", "Bond", "Angle"):
        assert getattr(mg.Energy, f) == pytest.approx(getattr(mo.Energy, f), rel=etol, abs=1e-300), f
    assert mg.Virial == pytest.approx(mo.Virial, rel=etol * 10)
    assert mg.BodyVirial == pytest.approx(mo.BodyVirial, rel=etol * 10, abs=1e-300)


def test_madelung_on_gpu(gpu):
    md, cfg = ionic.nacl_system(gpu, ncell=4, rc=3.5, accuracy=1e-6)
    assert md.Energy.Coulomb == pytest.approx(-cfg["N"] / 2 * ionic.MADELUNG_NACL, rel=1e-6)


@pytest.mark.parametrize("ncell,rc,acc", [(3, 2.9, 1e-6), (6, 4.0, 1e-5), (8, 5.0, 1e-7)])
def test_rock_salt(gpu, oracle, ncell, rc, acc):
    kw = dict(ncell=ncell, rc=rc, skin=0.2, accuracy=acc, jitter=0.2, kcoul=1.3)
    mg, cfg = ionic.nacl_system(gpu, **kw)
    mo, _ = ionic.nacl_system(oracle, **kw)
    compare(gpu, oracle, mg, mo, cfg["N"])


def test_coul_long_water(gpu, oracle):
    """test/test_coul_long.f90: rigid SPC/E, coul_long, kspace_ewald(1e-4)."""
    mg, cfg = water.spce_fixture_system(gpu, coul="long", kspace=1e-4)
    mo, _ = water.spce_fixture_system(oracle, coul="long", kspace=1e-4)
    compare(gpu, oracle, mg, mo, cfg["N"])
    for lib, md in ((gpu, mg), (oracle, mo)):
        lib.random_momenta(md, water.KB * 298.0, True, 86245)
        water.run(lib, md, 10, 2.0)
    H = lambda md: md.Energy.Potential + md.Kinetic.Total
    assert H(mg) == pytest.approx(H(mo), rel=1e-9)
    Rg, Ro = gpu.download(mg, "coordinates", (cfg["N"], 3)), oracle.download(mo, "coordinates", (cfg["N"], 3))
    assert np.max(np.abs(Rg - Ro)) < 1e-9


def test_virial_only_mode(gpu, oracle):
    """With Compute off the pair Coulomb energy is not summed, and W(long) =
    E(coul) + E(long) - W(coul) (EmDeeCode.f90:1251) follows it."""
    mg, cfg = water.spce_fixture_system(gpu, coul="long", kspace=1e-4)
    mo, _ = water.spce_fixture_system(oracle, coul="long", kspace=1e-4)
    for lib, md in ((gpu, mg), (oracle, mo)):
        md.Options.Compute = False
        lib.compute_forces(md)
    assert mg.Virial == pytest.approx(mo.Virial, rel=1e-11)


@pytest.mark.parametrize("coul", ["sf", "long"])
def test_flexible_water(gpu, oracle, coul):
    mg, cfg = ionic.flexible_water(gpu, coul=coul)
    mo, _ = ionic.flexible_water(oracle, coul=coul)
    compare(gpu, oracle, mg, mo, cfg["N"])
    assert mg.Energy.Bond > 0 and mg.Energy.Angle > 0
    for lib, md in ((gpu, mg), (oracle, mo)):
        lib.random_momenta(md, water.KB * 298.0, True, 1234)
        water.run(lib, md, 10, 0.5)
    H = lambda md: md.Energy.Potential + md.Kinetic.Total
    assert H(mg) == pytest.approx(H(mo), rel=1e-9)
    Rg, Ro = gpu.download(mg, "coordinates", (cfg["N"], 3)), oracle.download(mo, "coordinates", (cfg["N"], 3))
    assert np.max(np.abs(Rg - Ro)) < 1e-9


def test_bonds_without_pairs(gpu, oracle):
    """Bonded terms alone (no pair model active anywhere)."""
    R = np.array([[5.0, 5.0, 5.0], [6.2, 5.3, 4.9], [6.5, 6.4, 5.4], [7.1, 7.0, 6.3]])
    out = []
    for lib in (gpu, oracle):
        md = lib.system(1, 1, 3.0, 0.5, 4, None, None, None)
        lib.EmDee_set_pair_model(md, 1, 1, lib.EmDee_pair_none(), 0.0)
        for i in (1, 2, 3):
            lib.EmDee_add_bond(md, i, i + 1, lib.EmDee_bond_harmonic(2.0, 1.1))
        lib.EmDee_add_angle(md, 1, 2, 3, lib.EmDee_angle_harmonic(0.7, 1.9))
        lib.EmDee_add_angle(md, 2, 3, 4, lib.EmDee_angle_harmonic(0.5, 2.0))
        lib.upload(md, "box", np.array([20.0]))
        lib.upload(md, "coordinates", R)
        out.append((md, lib.download(md, "forces", (4, 3))))
    (mg, Fg), (mo, Fo) = out
    assert np.max(np.abs(Fg - Fo)) <= 1e-13 * np.max(np.abs(Fo))
    assert mg.Energy.Bond == pytest.approx(mo.Energy.Bond, rel=1e-14)
    assert mg.Energy.Angle == pytest.approx(mo.Energy.Angle, rel=1e-13)
    assert mg.Virial == pytest.approx(mo.Virial, rel=1e-13)
