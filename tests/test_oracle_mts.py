"""Multiple-time-step integration over model layers (emdee_amd.mts), on the
oracle: composed from EmDee_boost/EmDee_displace/EmDee_switch_model_layer
only (EmDeeCode.f90:929-946, 1024-1103), so the oracle runs the same calls.

Pins: with one inner step the r-RESPA step is velocity Verlet on the full
layer (the split F = (F1 - F2) + F2 only changes rounding); with four inner
steps the energy fluctuation at a 4 fs outer step drops to that of 1 fs
velocity Verlet, while the one-inner-step scheme at 4 fs fluctuates about
ten times more."""
import numpy as np
import pytest

from emdee_amd import mts
import water


def fresh(lib):
    md, cfg = water.two_layer_fixture(lib, False)
    lib.random_momenta(md, water.KB * 298.0, True, 7)
    return md, cfg["N"]


def test_one_inner_step_is_velocity_verlet(oracle):
    md, N = fresh(oracle)
    H = mts.respa(oracle, md, 4, 1.0, 1)
    R = oracle.download(md, "coordinates", (N, 3))
    mv, _ = fresh(oracle)
    oracle.switch_model_layer(mv, 1)
    Hv = []
    for _ in range(4):
        water.run(oracle, mv, 1, 1.0)
        Hv.append(mv.Energy.Potential + mv.Kinetic.Total)
    assert np.max(np.abs(R - oracle.download(mv, "coordinates", (N, 3)))) < 1e-9
    assert H == pytest.approx(Hv, rel=1e-12)


def test_inner_steps_restore_conservation(oracle):
    md, _ = fresh(oracle)
    h4 = np.std(mts.respa(oracle, md, 12, 4.0, 4))
    md, _ = fresh(oracle)
    h1 = np.std(mts.respa(oracle, md, 12, 4.0, 1))
    md, _ = fresh(oracle)
    oracle.switch_model_layer(md, 1)
    H = []
    for _ in range(48):
        water.run(oracle, md, 1, 1.0)
        H.append(md.Energy.Potential + md.Kinetic.Total)
    hv = np.std(H)
    assert h4 < 0.25 * h1
    assert h4 < 2.0 * hv
