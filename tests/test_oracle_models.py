"""Closed-form pins of the oracle's per-pair arithmetic.

A two-atom system at distance r (large box) gives E(r) = Energy.Potential,
W(r) = Virial and F = W r_vec / r^2.  Each model is checked against an
independent statement of it: the textbook formula, W = -r dE/dr by central
finite differences of E, and the defining properties of each modifier
(shift -> E(rc) = 0, shifted force -> F(rc) = 0 too, smoothing -> G(Rm) = 1,
G(rc) = 0 with zero slope).  uerfc is pinned by the published bound of
Abramowitz & Stegun 7.1.26, |error| <= 1.5e-7, against math.erfc.
"""
import ctypes
import math

import numpy as np
import pytest

L, RC, SKIN = 40.0, 3.0, 0.2


def pair_energy(lib, model_ctor, r, coul=None, q=(0.0, 0.0), kcoul=0.0, compute=True):
    R = np.array([[1.0, 2.0, 3.0], [1.0 + r, 2.0, 3.0]])
    md = lib.system(1, 1, RC, SKIN, 2, None, None, None)
    md.Options.Compute = compute
    lib.EmDee_set_pair_model(md, 1, 1, model_ctor(lib), kcoul)
    if coul is not None:
        lib.EmDee_set_coul_model(md, coul(lib))
    lib.upload(md, "charges", np.array(q))
    lib.upload(md, "box", np.array([L]))
    lib.upload(md, "coordinates", R)
    F = lib.download(md, "forces", (2, 3))
    return md.Energy.Potential, md.Virial, F


def fd_virial(lib, ctor, r, h=1e-6, **kw):
    ep = pair_energy(lib, ctor, r + h, **kw)[0]
    em = pair_energy(lib, ctor, r - h, **kw)[0]
    return -r * (ep - em) / (2 * h)


def lj_textbook(r, e=1.0, s=1.0):
    return 4 * e * ((s / r) ** 12 - (s / r) ** 6)


@pytest.mark.parametrize("r", [0.95, 1.12, 1.5, 2.2, 2.95])
def test_lj(oracle, r):
    E, W, F = pair_energy(oracle, lambda lib: lib.EmDee_pair_lj_cut(1.3, 1.1), r)
    assert E == pytest.approx(lj_textbook(r, 1.3, 1.1), rel=1e-13)
    assert W == pytest.approx(fd_virial(oracle, lambda lib: lib.EmDee_pair_lj_cut(1.3, 1.1), r), rel=1e-6)
    # F on atom 1 = W/r^2 * (R1 - R2): R1 - R2 = (-r, 0, 0)
    assert F[0, 0] == pytest.approx(-W / r, rel=1e-13)
    assert F[0] == pytest.approx(-F[1], rel=1e-15)


def test_beyond_cutoff_and_skin(oracle):
    E, W, F = pair_energy(oracle, lambda lib: lib.EmDee_pair_lj_cut(1.0, 1.0), RC + 0.1)
    assert E == 0.0 and W == 0.0 and np.all(F == 0.0)


@pytest.mark.parametrize("lam", [0.0, 0.3, 1.0])
@pytest.mark.parametrize("r", [0.5, 1.0, 1.7])
def test_softcore(oracle, lam, r):
    """Beutler soft-core LJ with alpha = 1/2, n = p = 1 (pair_softcore_cut.f90:30-32)."""
    ctor = lambda lib: lib.EmDee_pair_softcore_cut(0.9, 1.05, lam)
    E, W, _ = pair_energy(oracle, ctor, r)
    x = 0.5 * (1 - lam) + (r / 1.05) ** 6
    assert E == pytest.approx(4 * 0.9 * lam * (1 / x ** 2 - 1 / x), rel=1e-12, abs=1e-13)
    if lam > 0:
        assert W == pytest.approx(fd_virial(oracle, ctor, r), rel=1e-6)


def test_modifiers(oracle):
    lj = lambda lib: lib.EmDee_pair_lj_cut(1.0, 1.0)
    # shifted: E(r) = E_lj(r) - E_lj(rc), W unchanged
    E, W, _ = pair_energy(oracle, lambda lib: lib.EmDee_shifted(lj(lib)), 1.3)
    assert E == pytest.approx(lj_textbook(1.3) - lj_textbook(RC), rel=1e-13)
    # shifted force: E and F vanish at rc (approached from below)
    sf = lambda lib: lib.EmDee_shifted_force(lj(lib))
    E, W, _ = pair_energy(oracle, sf, RC - 1e-7)
    assert abs(E) < 1e-9 and abs(W) < 1e-7
    for r in (1.1, 2.0, 2.9):
        assert pair_energy(oracle, sf, r)[1] == pytest.approx(fd_virial(oracle, sf, r), rel=1e-6)
    # smoothed in r and in r^2: E = E_lj inside Rm, continuous, zero at rc
    for ctor in (lambda lib: lib.EmDee_smoothed(lj(lib), 1.0), lambda lib: lib.EmDee_square_smoothed(lj(lib), 1.0)):
        assert pair_energy(oracle, ctor, 1.5)[0] == pytest.approx(lj_textbook(1.5), rel=1e-13)
        assert abs(pair_energy(oracle, ctor, RC - 1e-6)[0]) < 1e-9
        Ein, Eout = pair_energy(oracle, ctor, 2.0 - 1e-7)[0], pair_energy(oracle, ctor, 2.0 + 1e-7)[0]
        assert Ein == pytest.approx(Eout, abs=1e-6)
        for r in (2.2, 2.6, 2.9):
            assert pair_energy(oracle, ctor, r)[1] == pytest.approx(fd_virial(oracle, ctor, r), rel=1e-6)
    # shifted smoothed: shift -(E(Rm) + E(rc))/2 applied everywhere, then switched
    for ctor in (lambda lib: lib.EmDee_shifted_smoothed(lj(lib), 1.0),
                 lambda lib: lib.EmDee_shifted_square_smoothed(lj(lib), 1.0)):
        shift = -0.5 * (lj_textbook(2.0) + lj_textbook(RC))
        assert pair_energy(oracle, ctor, 1.5)[0] == pytest.approx(lj_textbook(1.5) + shift, rel=1e-13)
        assert abs(pair_energy(oracle, ctor, RC - 1e-6)[0]) < 1e-9
        for r in (2.2, 2.6):
            assert pair_energy(oracle, ctor, r)[1] == pytest.approx(fd_virial(oracle, ctor, r), rel=1e-6)


def test_uerfc_published_bound(oracle):
    f = oracle.lib.oracle_uerfc
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_double]
    xs = np.linspace(0.0, 6.0, 2001)
    err = max(abs(f(x) - math.erfc(x)) for x in xs)
    assert err <= 1.5e-7
    assert err > 1e-9      # it is the approximation, not libm erfc


QQ = (0.7, -0.4)
KC = 1.3


def coul_only(lib, coul, r, compute=True):
    E, W, F = pair_energy(lib, lambda l: l.EmDee_pair_none(), r, coul=coul, q=QQ, kcoul=KC, compute=compute)
    return E, W


@pytest.mark.parametrize("r", [0.8, 1.7, 2.6])
def test_coulomb_models(oracle, r):
    qq = KC * QQ[0] * QQ[1]
    erfc = lambda x: oracle_uerfc_val(oracle, x)
    alpha = 0.4
    cases = {
        "cut": (lambda lib: lib.EmDee_coul_cut(), qq / r),
        # coul_sf: its shift is zeroed by modifier_setup after cutoff_setup
        # (EmDeeCode.f90:466-468), so it acts as the plain cut model.
        "sf": (lambda lib: lib.EmDee_coul_sf(), qq / r),
        "damped": (lambda lib: lib.EmDee_coul_damped(alpha), qq * erfc(alpha * r) / r),
        "shifted(damped)": (lambda lib: lib.EmDee_shifted(lib.EmDee_coul_damped(alpha)),
                            qq * (erfc(alpha * r) / r - erfc(alpha * RC) / RC)),
    }
    for name, (ctor, expect) in cases.items():
        E, W = coul_only(oracle, ctor, r)
        assert E == pytest.approx(expect, rel=1e-12), name
        ep = coul_only(oracle, ctor, r + 1e-6)[0]
        em = coul_only(oracle, ctor, r - 1e-6)[0]
        assert W == pytest.approx(-r * (ep - em) / 2e-6, rel=2e-6), name


def oracle_uerfc_val(oracle, x):
    f = oracle.lib.oracle_uerfc
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_double]
    return f(x)


def test_smoothed_coulomb_switches(oracle):
    """Square-smoothed models vanish at rc and are continuous at Rm;
    coul_damped_smoothed's switch is inert after set_coul_model (Rm reset to
    rc by modifier_setup), so it equals coul_damped."""
    for ctor in (lambda lib: lib.EmDee_coul_square_smoothed(1.0),
                 lambda lib: lib.EmDee_coul_damped_square_smoothed(0.4, 1.0),
                 lambda lib: lib.EmDee_coul_shifted_square_smoothed(1.0)):
        assert abs(coul_only(oracle, ctor, RC - 1e-7)[0]) < 1e-6
        a, b = coul_only(oracle, ctor, 2.0 - 1e-8)[0], coul_only(oracle, ctor, 2.0 + 1e-8)[0]
        assert a == pytest.approx(b, abs=1e-6)
        for r in (2.3, 2.8):
            ep, em = coul_only(oracle, ctor, r + 1e-6)[0], coul_only(oracle, ctor, r - 1e-6)[0]
            assert coul_only(oracle, ctor, r)[1] == pytest.approx(-r * (ep - em) / 2e-6, rel=2e-6)
    ds = coul_only(oracle, lambda lib: lib.EmDee_coul_damped_smoothed(0.4, 1.0), 2.7)
    dd = coul_only(oracle, lambda lib: lib.EmDee_coul_damped(0.4), 2.7)
    assert ds == pytest.approx(dd, rel=1e-15)


def test_virial_only_mode_matches(oracle):
    """Forces do not depend on Options.Compute (compute.f90 energy/virial
    variants), except for the documented algebraic variants."""
    for ctor in (lambda lib: lib.EmDee_shifted_force(lib.EmDee_pair_lj_cut(1, 1)),
                 lambda lib: lib.EmDee_square_smoothed(lib.EmDee_pair_lj_cut(1, 1), 1.0)):
        a = pair_energy(oracle, ctor, 2.4)[1]
        b = pair_energy(oracle, ctor, 2.4, compute=False)[1]
        assert a == pytest.approx(b, rel=1e-14)


def test_normal_generator_statistics(oracle):
    f = oracle.lib.oracle_normals
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    n = 400000
    x = np.zeros(n)
    f(86245, n, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert abs(x.mean()) < 5 / math.sqrt(n)
    assert abs(x.var() - 1) < 0.01
    assert abs(np.mean(x ** 4) - 3) < 0.05
    # tail mass beyond the ziggurat base strip r = 3.44262 (exercises the iz == 0 branch)
    assert abs(np.mean(np.abs(x) > 3.44262) - math.erfc(3.44262 / math.sqrt(2))) < 2e-4
