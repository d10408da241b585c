"""Phase-space sharing on the oracle (EmDeeCode.f90:239-269): after
EmDee_share_phase_space the losing system works on the keeper's coordinates,
momenta and bodies, so its energy equals that of a fresh system built from
the shared configuration, and dynamics on either system moves both."""
import numpy as np
import pytest

import water


def test_shared_configuration_energy(oracle):
    md, cfg = water.spce_fixture_system(oracle, rc=10.0, skin=2.0)
    other, _ = water.spce_fixture_system(oracle, rc=5.0, skin=2.0)
    N = cfg["N"]
    oracle.random_momenta(md, water.KB * 298.0, True, 5)
    water.run(oracle, md, 5, 1.0)
    oracle.EmDee_share_phase_space(md, other)
    water.run(oracle, md, 5, 1.0)
    other.Options.Compute = True
    oracle.compute_forces(other)
    R = oracle.download(md, "coordinates", (N, 3))
    fresh, _ = water.spce_fixture_system(oracle, rc=5.0, skin=2.0)
    oracle.upload(fresh, "coordinates", R)
    oracle.compute_forces(fresh)
    assert other.Energy.Potential == pytest.approx(fresh.Energy.Potential, rel=1e-10)
    assert np.array_equal(oracle.download(other, "coordinates", (N, 3)), R)
    water.run(oracle, other, 3, 1.0)
    assert np.array_equal(oracle.download(md, "coordinates", (N, 3)), oracle.download(other, "coordinates", (N, 3)))
