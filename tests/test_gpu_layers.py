"""Model layers on the GPU against the oracle (EmDeeCode.f90:273-305,
363-412, 486-520, 929-946): two layers with different Coulomb models, an
inner cutoff on one of them (layer_based_parameters), layer switching with
per-layer forces, energies and virials, and dynamics on either layer."""
import numpy as np
import pytest

from emdee_amd import mts
from emdee_amd import surrogate as sg
import water

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("inner", [False, True])
def test_two_layers(gpu, oracle, inner):
    (mg, cfg), (mo, _) = water.two_layer_fixture(gpu, inner), water.two_layer_fixture(oracle, inner)
    N = cfg["N"]
    for layer in (1, 2, 1):
        for lib, md in ((gpu, mg), (oracle, mo)):
            lib.switch_model_layer(md, layer)
        Fg, Fo = gpu.download(mg, "forces", (N, 3)), oracle.download(mo, "forces", (N, 3))
        assert np.max(np.abs(Fg - Fo)) <= 1e-11 * np.max(np.abs(Fo)), layer
        assert mg.Energy.Potential == pytest.approx(mo.Energy.Potential, rel=1e-12)
        assert mg.Virial == pytest.approx(mo.Virial, rel=1e-11)
    # a few steps on layer 2, then the layer-1 view of the new state
    for lib, md in ((gpu, mg), (oracle, mo)):
        lib.random_momenta(md, water.KB * 298.0, True, 7)
        lib.switch_model_layer(md, 2)
        water.run(lib, md, 5, 1.0)
        lib.switch_model_layer(md, 1)
        lib.compute_forces(md)
    assert mg.Energy.Potential == pytest.approx(mo.Energy.Potential, rel=1e-10)
    assert mg.Builds == mo.Builds


def test_atom_energies_sum_to_the_potential(gpu):
    """EmDeeGPU_atom_energies: atom energies add up to the layer's potential
    energy less the constant intrabody k-space term (rigid water)."""
    md, cfg = sg.two_layer_water(gpu, 256, seed=2)
    N = cfg["N"]
    for layer in (1, 2):
        gpu.switch_model_layer(md, layer)
        e = gpu.atom_energies(md, N)
        gpu.compute_forces(md)
        if layer == 2:
            assert e.sum() == pytest.approx(md.Energy.Potential, rel=1e-12)
        else:
            e_layer1 = e.sum()
    # layer 1 differs by the intrabody discount only: the same for any configuration
    gpu.switch_model_layer(md, 1)
    gpu.compute_forces(md)
    c1 = md.Energy.Potential - e_layer1
    gpu.random_momenta(md, sg.KB * 298.0, True, 3)
    sg.trajectory(gpu, md, 1, 5, 1.0)
    e = gpu.atom_energies(md, N)
    gpu.compute_forces(md)
    assert md.Energy.Potential - e.sum() == pytest.approx(c1, rel=1e-9)


def test_respa_matches_oracle(gpu, oracle):
    """r-RESPA over the two layers (emdee_amd.mts, 1:4) on the GPU against
    the same reference-API calls on the oracle."""
    runs = []
    for lib in (gpu, oracle):
        md, cfg = water.two_layer_fixture(lib, False)
        lib.random_momenta(md, water.KB * 298.0, True, 7)
        H = mts.respa(lib, md, 4, 2.0, 4)
        runs.append((H, lib.download(md, "coordinates", (cfg["N"], 3)), md.Builds))
    (Hg, Rg, bg), (Ho, Ro, bo) = runs
    assert Hg == pytest.approx(Ho, rel=1e-10)
    assert np.max(np.abs(Rg - Ro)) < 1e-8
    assert bg == bo
