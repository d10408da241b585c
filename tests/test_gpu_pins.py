"""The HIP engine reproduces the energy pins of the reference's own tests.

test/test_pair_lj_cut.f90:46 and test/test_pair_lj_sf.f90:46 hold [U, W, U+K]
after 100 steps of NIST_lj_sample.mol at absolute tolerance 1e-8.  The
square-smoothed pin (test/test_pair_lj_smoothed.f90:46) was produced when the
modifier's second argument meant the switching radius: it is reproduced with
skin width rc - Rm = 1.0.
"""
import numpy as np
import pytest

import hosts

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["lj_cut", "lj_sf"])
def test_reference_pin(gpu, name):
    pins = hosts.load_pins()
    md, _ = hosts.lj_sample_system(gpu, name)
    got = hosts.run_md(gpu, md, 100, 0.002, 20)
    print(name, got, got - pins["pins"][name]["expected"], "builds", md.Builds, flush=True)
    assert np.all(np.abs(got - np.array(pins["pins"][name]["expected"])) <= pins["tolerance"])


def test_square_smoothed_pin(gpu):
    pins = hosts.load_pins()
    cfg = hosts.read_configuration(hosts.GOLDEN + "/NIST_lj_sample.mol")
    md = gpu.system(1, 1, 3.0, 0.3, cfg["N"], cfg["types"], cfg["mass"], None)
    gpu.EmDee_set_pair_model(md, 1, 1, gpu.EmDee_square_smoothed(gpu.EmDee_pair_lj_cut(1.0, 1.0), 1.0), 0.0)
    gpu.upload(md, "charges", cfg["Q"])
    gpu.upload(md, "box", np.array([cfg["L"]]))
    gpu.upload(md, "coordinates", cfg["R"])
    gpu.random_momenta(md, 0.8, True, 86245)
    got = hosts.run_md(gpu, md, 100, 0.002, 20)
    assert np.all(np.abs(got - np.array(pins["pins"]["lj_smoothed"]["expected"])) <= pins["tolerance"])
