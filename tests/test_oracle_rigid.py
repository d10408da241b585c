"""Pins of the oracle's rigid-body restatement (ArBee.f90) by mathematics:

* principal moments from Kopp's analytic eigensolver equal numpy.linalg.eigh
  of the inertia tensor built independently from the fixture coordinates;
* quaternions stay normalised; body-frame coordinates are preserved;
* exact torque-free rotation conserves rotational kinetic energy and the
  space-frame angular momentum of an asymmetric top to round-off;
* the exact rotation and NO_SQUISH with many substeps converge to each other;
* particle momenta of a body sum to its centre-of-mass momentum.
"""
import numpy as np
import pytest

import hosts
import water


def quat_matrix(q):
    """Rotation matrix A (space -> body) of a unit quaternion (Shepperd's
    convention, math.f90:189-218 inverted)."""
    q0, q1, q2, q3 = q
    return np.array([
        [q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3, 2 * (q1 * q2 + q0 * q3), 2 * (q1 * q3 - q0 * q2)],
        [2 * (q1 * q2 - q0 * q3), q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3, 2 * (q2 * q3 + q0 * q1)],
        [2 * (q1 * q3 + q0 * q2), 2 * (q2 * q3 - q0 * q1), q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3]])


def test_inertia_against_eigh(oracle):
    md, cfg = water.spce_fixture_system(oracle)
    nb = cfg["N"] // 3
    MoI = oracle.download(md, "inertia", (nb, 3))
    R, L = cfg["R"], cfg["L"]
    m = cfg["mass"][cfg["types"] - 1]
    for b in range(0, nb, 37):
        idx = np.arange(3 * b, 3 * b + 3)
        X = R[idx].copy()
        X[1:] -= L * np.round((X[1:] - X[0]) / L)
        cm = (m[idx, None] * X).sum(0) / m[idx].sum()
        d = X - cm
        I = sum(m[k] * (np.dot(d[j], d[j]) * np.eye(3) - np.outer(d[j], d[j])) for j, k in enumerate(idx))
        w = np.linalg.eigh(I)[0]
        assert sorted(np.abs(MoI[b])) == pytest.approx(sorted(np.abs(w)), rel=1e-10)
        # ordering: decreasing magnitude (math.f90:296-299)
        assert abs(MoI[b, 0]) >= abs(MoI[b, 1]) >= abs(MoI[b, 2])


def test_quaternion_frame(oracle):
    md, cfg = water.spce_fixture_system(oracle)
    nb = cfg["N"] // 3
    q = oracle.download(md, "quaternions", (nb, 4))
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-14)
    # space-frame offsets rotated by A(q) are diagonal-inertia body coordinates
    R = oracle.download(md, "coordinates", (cfg["N"], 3))
    cm = oracle.download(md, "centersOfMass", (nb, 3))
    MoI = oracle.download(md, "inertia", (nb, 3))
    m = cfg["mass"][cfg["types"] - 1]
    for b in range(0, nb, 53):
        idx = np.arange(3 * b, 3 * b + 3)
        dbody = (quat_matrix(q[b]) @ (R[idx] - cm[b]).T).T
        I = sum(m[k] * (np.dot(dbody[j], dbody[j]) * np.eye(3) - np.outer(dbody[j], dbody[j])) for j, k in enumerate(idx))
        assert np.allclose(I, np.diag(MoI[b]), atol=1e-9 * np.max(np.abs(MoI[b])))


def single_top(lib, mode):
    """One asymmetric rigid molecule, no interactions: free rotation."""
    R = np.array([[0.0, 0.0, 0.0], [1.0, 0.2, 0.1], [-0.3, 0.9, -0.2], [0.4, -0.1, 0.8]]) + 5.0
    md = lib.system(1, 1, 2.0, 0.5, 4, None, np.array([1.0]), np.array([1, 1, 1, 1], dtype=np.int32))
    md.Options.RotationMode = mode
    lib.upload(md, "box", np.array([30.0]))
    lib.upload(md, "coordinates", R)
    P = np.array([[0.3, -0.2, 0.5], [-0.7, 0.4, 0.1], [0.2, 0.9, -0.6], [0.1, -0.3, 0.2]])
    lib.upload(md, "momenta", P)
    return md


def space_angular_momentum(lib, md):
    R = lib.download(md, "coordinates", (4, 3))
    P = lib.download(md, "momenta", (4, 3))
    cm = R.mean(0)
    return np.sum(np.cross(R - cm, P), axis=0)


def test_exact_free_rotation_conserves(oracle):
    md = single_top(oracle, 0)
    K0 = md.Kinetic.Rotational
    L0 = space_angular_momentum(oracle, md)
    for _ in range(400):
        oracle.boost(md, 1.0, 0.0, 0.05)
        oracle.displace(md, 1.0, 0.0, 0.1)
        oracle.boost(md, 1.0, 0.0, 0.05)
    assert md.Kinetic.Rotational == pytest.approx(K0, rel=1e-11)
    assert space_angular_momentum(oracle, md) == pytest.approx(L0, rel=1e-10, abs=1e-12)
    q = oracle.download(md, "quaternions", (1, 4))
    assert np.linalg.norm(q) == pytest.approx(1.0, abs=1e-14)


def test_exact_matches_no_squish_limit(oracle):
    a, b = single_top(oracle, 0), single_top(oracle, 200)
    for md in (a, b):
        oracle.displace(md, 1.0, 0.0, 0.7)
    qa = oracle.download(a, "quaternions", (1, 4))[0]
    qb = oracle.download(b, "quaternions", (1, 4))[0]
    assert min(np.linalg.norm(qa - qb), np.linalg.norm(qa + qb)) < 1e-7
    Ra, Rb = oracle.download(a, "coordinates", (4, 3)), oracle.download(b, "coordinates", (4, 3))
    assert np.max(np.abs(Ra - Rb)) < 1e-7


def test_body_momenta_consistency(oracle):
    md, cfg = water.spce_fixture_system(oracle)
    oracle.random_momenta(md, water.KB * 298.0, True, 86245)
    nb = cfg["N"] // 3
    P = oracle.download(md, "momenta", (cfg["N"], 3))
    pcm = oracle.download(md, "bodymom", (nb, 3))
    assert np.allclose(P.reshape(nb, 3, 3).sum(1), pcm, rtol=1e-12, atol=1e-12 * np.max(np.abs(pcm)))
    # equipartition after the exact rescale: 2K = (6 nb - 3) kT
    assert 2 * md.Kinetic.Total == pytest.approx((6 * nb - 3) * water.KB * 298.0, rel=1e-12)
