"""Pins of the oracle's dihedral term (EmDeeData.f90:554-640 restated, with
the harmonic model E = k/2 (phi - phi0)^2, difference in (-pi, pi]):

* phi is the IUPAC dihedral angle of i-j-k-l, computed here independently
  (atan2(|b2| b1 . (b2 x b3), (b1 x b2) . (b2 x b3))), so the energy has a
  closed form;
* the forces are minus the gradient of that energy (central differences),
  including across the +-pi wrap of phi - phi0;
* the four forces sum to zero and the virial of an isolated dihedral is zero
  (its energy does not change when all positions scale).
A sign slip in the restated force expressions, or the reference's own
+(Fi + Fk + Fl) on atom j, fails the sum and the difference checks."""
import numpy as np
import pytest


def iupac(R):
    b1, b2, b3 = R[1] - R[0], R[2] - R[1], R[3] - R[2]
    return np.arctan2(np.linalg.norm(b2) * np.dot(b1, np.cross(b2, b3)), np.dot(np.cross(b1, b2), np.cross(b2, b3)))


def chain(lib, R, k, phi0, L=20.0):
    md = lib.system(1, 1, 3.0, 0.5, len(R), None, None, None)
    lib.EmDee_set_pair_model(md, 1, 1, lib.EmDee_pair_none(), 0.0)
    for c in range(len(R) // 4):
        lib.EmDee_add_dihedral(md, 4 * c + 1, 4 * c + 2, 4 * c + 3, 4 * c + 4, lib.EmDee_dihedral_harmonic(k, phi0))
    lib.upload(md, "box", np.array([L]))
    lib.upload(md, "coordinates", R)
    return md


def wrap(d):
    return d - 2 * np.pi * np.round(d / (2 * np.pi))


@pytest.mark.parametrize("seed,phi0", [(1, 0.7), (2, -2.9), (3, 3.1), (4, 0.0)])
def test_dihedral_energy_and_forces(oracle, seed, phi0):
    rng = np.random.default_rng(seed)
    base = np.array([[0, 0, 0], [1.5, 0, 0], [2.0, 1.4, 0], [3.4, 1.6, 0.9]], float)
    R = base + 0.3 * rng.standard_normal((4, 3)) + 8.0
    k = 2.0
    md = chain(oracle, R, k, phi0)
    F = oracle.download(md, "forces", (4, 3))
    assert md.Energy.Dihedral == pytest.approx(0.5 * k * wrap(iupac(R) - phi0) ** 2, rel=1e-12, abs=1e-14)
    assert md.Energy.Potential == pytest.approx(md.Energy.Dihedral, rel=1e-14)
    assert np.max(np.abs(F.sum(0))) < 1e-12 * np.max(np.abs(F))
    assert abs(md.Virial) < 1e-12 * np.max(np.abs(F))
    h = 1e-6
    for a in range(4):
        for x in range(3):
            Rp, Rm = R.copy(), R.copy()
            Rp[a, x] += h
            Rm[a, x] -= h
            d = chain(oracle, Rp, k, phi0).Energy.Potential - chain(oracle, Rm, k, phi0).Energy.Potential
            assert F[a, x] == pytest.approx(-d / (2 * h), rel=1e-6, abs=1e-8)


def test_dihedrals_across_the_box_boundary(oracle):
    """A chain split by the periodic boundary (minimum image on each bond
    vector) has the energy and forces of the same chain kept whole."""
    R = np.array([[0.2, 0.3, 0.1], [1.6, 0.1, 0.3], [2.1, 1.5, 0.0], [3.5, 1.7, 1.0]])
    whole = chain(oracle, R + 5.0, 1.5, 1.0, L=12.0)
    split = chain(oracle, R - 1.0, 1.5, 1.0, L=12.0)
    assert split.Energy.Dihedral == pytest.approx(whole.Energy.Dihedral, rel=1e-12)
    assert np.allclose(oracle.download(split, "forces", (4, 3)), oracle.download(whole, "forces", (4, 3)),
                       rtol=1e-10, atol=1e-12)
