"""The oracle reproduces the energy pins held by the reference's own tests
(test/test_pair_lj_*.f90:46): [U, W, U+K] after 100 steps of NIST_lj_sample
at absolute tolerance 1e-8 (test/common/declarations.f90:16).  These pins fix
the RNG stream, momentum adjustment, integrator, neighbor lists and the LJ /
shifted-force pair arithmetic together."""
import numpy as np
import pytest

import hosts


@pytest.mark.parametrize("threads", [1, 2])
@pytest.mark.parametrize("name", ["lj_cut", "lj_sf"])
def test_reference_pins(oracle, name, threads):
    pins = hosts.load_pins()
    md, _ = hosts.lj_sample_system(oracle, name, threads=threads)
    got = hosts.run_md(oracle, md, 100, 0.002, 20)
    assert np.all(np.abs(got - np.array(pins["pins"][name]["expected"])) <= pins["tolerance"])


def smoothed_run(lib, ctor, skin):
    cfg = hosts.read_configuration(hosts.GOLDEN + "/NIST_lj_sample.mol")
    md = lib.system(2, 1, 3.0, 0.3, cfg["N"], cfg["types"], cfg["mass"], None)
    lib.EmDee_set_pair_model(md, 1, 1, ctor(lib.EmDee_pair_lj_cut(1.0, 1.0), skin), 0.0)
    lib.upload(md, "charges", cfg["Q"])
    lib.upload(md, "box", np.array([cfg["L"]]))
    lib.upload(md, "coordinates", cfg["R"])
    lib.random_momenta(md, 0.8, True, 86245)
    return hosts.run_md(lib, md, 100, 0.002, 20)


def test_square_smoothed_pin_with_switch_width(oracle):
    """test_pair_lj_smoothed.f90:46 was produced when the second argument was
    the switching radius (the test passes Rc - 1); with the snapshot's skin
    semantics that is skin = rc - Rm = 1.0."""
    pins = hosts.load_pins()
    got = smoothed_run(oracle, oracle.EmDee_square_smoothed, 1.0)
    assert np.all(np.abs(got - np.array(pins["pins"]["lj_smoothed"]["expected"])) <= pins["tolerance"])
    # as literally written (skin = Rc - 1 = 2) the pin does not hold
    bad = smoothed_run(oracle, oracle.EmDee_square_smoothed, 2.0)
    assert np.max(np.abs(bad - np.array(pins["pins"]["lj_smoothed"]["expected"]))) > 1.0


def test_shifted_square_smoothed_pin_under_earlier_shift(oracle):
    """test_pair_lj_shifted_smoothed.f90:46 was produced by an earlier library
    than the src/ snapshot.  Its trajectory and virial are the snapshot's
    square-smoothed arithmetic with switch width 1.0 (as for the smoothed pin)
    and the energy shift -E(rc); its energy adds that shift only inside the
    switching region.  Under that rule the checker reproduces all three
    numbers; under the snapshot's rule (shift -(E(Rm) + E(rc))/2 at every
    distance, modelClass_nonbonded.f90:281-290) it cannot, and the gap is
    large, so the pin discriminates between the two rules."""
    pins = hosts.load_pins()
    expected = np.array(pins["pins"]["lj_shifted_smoothed"]["expected"])
    oracle.lib.oracle_set_legacy_shifted_smoothing(1)
    try:
        got = smoothed_run(oracle, oracle.EmDee_shifted_square_smoothed, 1.0)
    finally:
        oracle.lib.oracle_set_legacy_shifted_smoothing(0)
    assert np.all(np.abs(got - expected) <= pins["tolerance"]), got - expected
    snapshot = smoothed_run(oracle, oracle.EmDee_shifted_square_smoothed, 1.0)
    # the snapshot's shift moves the forces inside the switching region
    assert np.max(np.abs(snapshot - expected)) > 100.0
