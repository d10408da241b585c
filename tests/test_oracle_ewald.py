"""Pins of the oracle's Ewald sum (kspace_ewald.f90, modelClass_kspace.f90,
coul_long.f90) and bonded terms (EmDeeData.f90:443-550) by textbook facts:

* the Madelung constant of rock salt, 1.747564594633...;
* an independent Ewald restatement in numpy (full k sphere instead of the
  reference's half space, complex exponentials instead of recursions, the
  Ewald parameters solved with scipy) for energy and forces;
* alpha independence of the total Coulomb energy;
* excluded pairs (rigid bodies, bonds, angles): the engine's energy equals
  the full Ewald energy of the same charges minus the bare Coulomb energy
  q_i q_j / r of every excluded pair (real-space term skipped, erf part
  discounted from k space);
* harmonic bond and angle energies in closed form, forces = -grad E by
  central differences.
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

import ionic
import water

A_S = (0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429, 0.327591100)


def uerfc_published(x):
    """Abramowitz & Stegun 7.1.26 (the approximation the reference uses)."""
    a1, a2, a3, a4, a5, p = A_S
    t = 1.0 / (1.0 + p * x)
    return t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5)))) * np.exp(-x * x)


def ewald_parameters(accuracy, rc, L):
    # s^2 + ln s^2 = -ln(accuracy), alpha = s / rc, kmax = 2 alpha s
    y = -math.log(accuracy)
    s2 = brentq(lambda u: u + math.log(u) - y, 1e-6, 1e3, xtol=1e-15)
    s = math.sqrt(s2)
    alpha = s / rc
    nmax = math.ceil(2 * alpha * s / (2 * math.pi / L))
    return alpha, nmax


def textbook_ewald(R, Q, L, rc, accuracy, kcoul=1.0):
    """Energy and forces of point charges by the Ewald method with the
    reference's choices of alpha, k sphere (|n| <= nmax) and real-space
    cutoff, written independently (Frenkel & Smit eq. 12.1.25-12.1.28)."""
    alpha, nmax = ewald_parameters(accuracy, rc, L)
    N = len(Q)
    # real space, minimum image, r < rc
    E_real = 0.0
    F = np.zeros((N, 3))
    for i in range(N - 1):
        d = R[i] - R[i + 1:]
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(1))
        m = r < rc
        qq = kcoul * Q[i] * Q[i + 1:][m]
        x = alpha * r[m]
        e = qq * uerfc_published(x) / r[m]
        w = e + qq * 2 * alpha / math.sqrt(math.pi) * np.exp(-x * x)
        E_real += e.sum()
        f = (w / r[m] ** 2)[:, None] * d[m]
        F[i] += f.sum(0)
        F[np.arange(i + 1, N)[m]] -= f
    # reciprocal space, full sphere |n|^2 <= nmax^2, n != 0
    g = np.arange(-nmax, nmax + 1)
    n = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n = n[((n * n).sum(1) <= nmax * nmax) & ((n != 0).any(1))]
    k = 2 * math.pi / L * n
    k2 = (k * k).sum(1)
    f_k = np.exp(-k2 / (4 * alpha * alpha)) / k2
    V = L ** 3
    phase = np.exp(1j * (R @ k.T))                # [N, nk]
    S = (Q[:, None] * phase).sum(0)
    E_recip = kcoul * 2 * math.pi / V * (f_k * np.abs(S) ** 2).sum()
    F += kcoul * 4 * math.pi / V * Q[:, None] * ((np.conj(S)[None, :] * phase).imag * f_k) @ k
    E_self = -kcoul * alpha / math.sqrt(math.pi) * (Q * Q).sum()
    return E_real + E_recip + E_self, F


@pytest.mark.parametrize("accuracy,tol", [(1e-5, 1e-5), (1e-6, 1e-6), (1e-8, 2e-6)])
def test_madelung_rock_salt(oracle, accuracy, tol):
    md, cfg = ionic.nacl_system(oracle, ncell=4, rc=3.5, accuracy=accuracy)
    expected = -cfg["N"] / 2 * ionic.MADELUNG_NACL
    assert md.Energy.Coulomb == pytest.approx(expected, rel=tol)
    # forces vanish by symmetry on the perfect lattice
    F = oracle.download(md, "forces", (cfg["N"], 3))
    assert np.max(np.abs(F)) < 1e-9


def test_against_textbook_ewald(oracle):
    md, cfg = ionic.nacl_system(oracle, ncell=3, rc=2.9, skin=0.05, accuracy=1e-6, jitter=0.15, kcoul=1.3)
    E, F = textbook_ewald(cfg["R"], cfg["Q"], cfg["L"], 2.9, 1e-6, kcoul=1.3)
    assert md.Energy.Coulomb == pytest.approx(E, rel=1e-11)
    Fo = oracle.download(md, "forces", (cfg["N"], 3))
    assert np.max(np.abs(Fo - F)) <= 1e-10 * np.max(np.abs(F))
    # Coulomb energy is homogeneous of degree -1: W = E
    assert md.Virial == pytest.approx(md.Energy.Coulomb, rel=1e-12)


def test_alpha_independence(oracle):
    Es = []
    for acc in (1e-7, 1e-9):
        md, cfg = ionic.nacl_system(oracle, ncell=3, rc=2.9, accuracy=acc, jitter=0.2)
        Es.append(md.Energy.Coulomb)
    assert Es[0] == pytest.approx(Es[1], rel=2e-6)


def bare_coulomb(R, Q, L, pairs, kcoul):
    e = 0.0
    for i, j in pairs:
        d = R[i] - R[j]
        d -= L * np.round(d / L)
        e += kcoul * Q[i] * Q[j] / np.sqrt(d @ d)
    return e


def charges_only_water(lib, bodies):
    cfg = ionic.hosts.read_configuration(ionic.hosts.GOLDEN + "/NIST_spce_sample.mol")
    md = lib.system(1, 1, 9.0, 1.0, cfg["N"], cfg["types"], cfg["mass"], cfg["molecule"] if bodies else None)
    for i in range(cfg["ntypes"]):
        lib.EmDee_set_pair_model(md, i + 1, i + 1, lib.EmDee_pair_none(), water.KCOUL)
    lib.EmDee_set_coul_model(md, lib.EmDee_coul_long())
    lib.EmDee_set_kspace_model(md, lib.EmDee_kspace_ewald(1e-6))
    return md, cfg


def finish(lib, md, cfg):
    lib.upload(md, "charges", cfg["Q"])
    lib.upload(md, "coordinates", cfg["R"])
    lib.upload(md, "box", np.array([cfg["L"]]))
    return md.Energy.Coulomb


def intramolecular_pairs(cfg):
    mol = cfg["molecule"]
    out = []
    for m in np.unique(mol):
        idx = np.where(mol == m)[0]
        out += [(a, b) for ai, a in enumerate(idx) for b in idx[ai + 1:]]
    return out


def test_rigid_body_exclusions(oracle):
    md, cfg = charges_only_water(oracle, bodies=False)
    E_all = finish(oracle, md, cfg)
    md, cfg = charges_only_water(oracle, bodies=True)
    E_rigid = finish(oracle, md, cfg)
    E_intra = bare_coulomb(cfg["R"], cfg["Q"], cfg["L"], intramolecular_pairs(cfg), water.KCOUL)
    assert E_rigid == pytest.approx(E_all - E_intra, rel=1e-11)


def test_bond_and_angle_exclusions(oracle):
    md, cfg = charges_only_water(oracle, bodies=False)
    E_all = finish(oracle, md, cfg)
    md, cfg = charges_only_water(oracle, bodies=False)
    types = cfg["types"]
    for m in np.unique(cfg["molecule"]):
        idx = np.where(cfg["molecule"] == m)[0] + 1
        o = [i for i in idx if types[i - 1] == 1][0]
        hs = [i for i in idx if types[i - 1] != 1]
        for h in hs:
            oracle.EmDee_add_bond(md, int(o), int(h), oracle.EmDee_bond_none())
        oracle.EmDee_add_angle(md, int(hs[0]), int(o), int(hs[1]), oracle.EmDee_angle_none())
    E_ex = finish(oracle, md, cfg)
    E_intra = bare_coulomb(cfg["R"], cfg["Q"], cfg["L"], intramolecular_pairs(cfg), water.KCOUL)
    assert E_ex == pytest.approx(E_all - E_intra, rel=1e-11)
    assert md.Energy.Bond == 0.0 and md.Energy.Angle == 0.0


def three_atoms(lib, R, kb=2.0, r0=1.1, ka=0.7, th0=1.9):
    md = lib.system(1, 1, 3.0, 0.5, 3, None, None, None)
    lib.EmDee_set_pair_model(md, 1, 1, lib.EmDee_pair_none(), 0.0)
    lib.EmDee_add_bond(md, 1, 2, lib.EmDee_bond_harmonic(kb, r0))
    lib.EmDee_add_bond(md, 2, 3, lib.EmDee_bond_harmonic(kb, r0))
    lib.EmDee_add_angle(md, 1, 2, 3, lib.EmDee_angle_harmonic(ka, th0))
    lib.upload(md, "box", np.array([20.0]))
    lib.upload(md, "coordinates", R)
    return md


def test_harmonic_bond_angle_closed_form(oracle):
    R = np.array([[5.0, 5.0, 5.0], [6.2, 5.3, 4.9], [6.5, 6.4, 5.4]])
    md = three_atoms(oracle, R)
    r12, r23 = np.linalg.norm(R[0] - R[1]), np.linalg.norm(R[2] - R[1])
    assert md.Energy.Bond == pytest.approx(0.5 * 2.0 * ((r12 - 1.1) ** 2 + (r23 - 1.1) ** 2), rel=1e-13)
    a, b = R[0] - R[1], R[2] - R[1]
    th = math.acos(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
    assert md.Energy.Angle == pytest.approx(0.5 * 0.7 * (th - 1.9) ** 2, rel=1e-12)
    F = oracle.download(md, "forces", (3, 3))
    h = 1e-6
    for i in range(3):
        for x in range(3):
            Rp, Rm = R.copy(), R.copy()
            Rp[i, x] += h
            Rm[i, x] -= h
            ep, em = three_atoms(oracle, Rp).Energy.Potential, three_atoms(oracle, Rm).Energy.Potential
            assert F[i, x] == pytest.approx(-(ep - em) / (2 * h), abs=1e-7)
    assert np.allclose(F.sum(0), 0.0, atol=1e-13)
    # virial W = sum r . F for a translation-invariant potential
    assert md.Virial == pytest.approx(np.sum(R * F), rel=1e-10)
