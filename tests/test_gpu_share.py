"""EmDee_share_phase_space on the GPU against the oracle
(EmDeeCode.f90:239-269; test/test_phase_space_sharing.f90): two rigid-water
systems with different cutoffs share coordinates, momenta, bodies and box;
dynamics driven through either one is seen by the other."""
import numpy as np
import pytest

import water

pytestmark = pytest.mark.gpu


def pair_of_systems(lib):
    md, cfg = water.spce_fixture_system(lib, rc=10.0, skin=2.0)
    other, _ = water.spce_fixture_system(lib, rc=5.0, skin=2.0)
    return md, other, cfg


def test_sharing_matches_oracle(gpu, oracle):
    runs = []
    for lib in (gpu, oracle):
        md, other, cfg = pair_of_systems(lib)
        N = cfg["N"]
        lib.random_momenta(md, water.KB * 298.0, True, 5)
        water.run(lib, md, 10, 1.0)
        lib.EmDee_share_phase_space(md, other)
        other.Options.Compute = True
        lib.compute_forces(other)
        rec = [(other.Energy.Potential, other.Virial, other.Kinetic.Total)]
        water.run(lib, md, 10, 1.0)            # drive the keeper
        lib.compute_forces(other)
        rec.append((other.Energy.Potential, other.Virial))
        water.run(lib, other, 10, 1.0)         # drive the other one
        lib.compute_forces(md)
        rec.append((md.Energy.Potential, md.Virial))
        Rm, Ro = lib.download(md, "coordinates", (N, 3)), lib.download(other, "coordinates", (N, 3))
        Pm, Po = lib.download(md, "momenta", (N, 3)), lib.download(other, "momenta", (N, 3))
        assert np.array_equal(Rm, Ro) and np.array_equal(Pm, Po)
        runs.append((rec, Rm))
    (rg, Rg), (ro, Ro) = runs
    for a, b in zip(rg, ro):
        assert a == pytest.approx(b, rel=1e-9)
    assert np.max(np.abs(Rg - Ro)) < 1e-9
