"""The oracle's neighbor lists (neighbor_lists.f90 restated) against a
brute-force all-pairs set, and its rebuild trigger against the closed form
of maximum_approach_sq (neighbor_lists.f90:41-59)."""
import ctypes

import numpy as np
import pytest

import parity_util as pu


def lj(lib):
    return lib.EmDee_pair_lj_cut(1.0, 1.0)


@pytest.mark.parametrize("threads", [1, 3])
@pytest.mark.parametrize("case", ["nist", "fcc", "small_box_M5", "random_outside_box"])
def test_oracle_pairs_brute_force(oracle, case, threads):
    import hosts
    if case == "nist":
        cfg = hosts.read_configuration(hosts.GOLDEN + "/NIST_lj_sample.mol")
        R, L, rc, skin = cfg["R"], cfg["L"], 3.0, 0.3
    elif case == "fcc":
        R, L = pu.fcc(1372, 0.8442, jitter=0.05, seed=1)
        rc, skin = 2.5, 0.3
    elif case == "small_box_M5":
        R, L = pu.fcc(256, 0.8442, jitter=0.1, seed=2)
        rc, skin = 2.5, 0.3
    else:
        L = 9.0
        R = pu.random_box(300, L, seed=3)
        rc, skin = 2.0, 0.4
    md = pu.build(oracle, R, L, rc, skin, [lj], threads=threads)
    got = pu.oracle_pairs(oracle, md)
    assert got == pu.brute_pairs(R, L, rc + skin)


def test_stencil_is_half_space(oracle):
    f = oracle.lib.oracle_stencil
    out = np.zeros(62 * 3, dtype=np.int32)
    f(out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    st = [tuple(x) for x in out.reshape(62, 3)]
    assert len(set(st)) == 62
    full = {(a, b, c) for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3)} - {(0, 0, 0)}
    assert set(st) | {(-a, -b, -c) for a, b, c in st} == full
    assert not (set(st) & {(-a, -b, -c) for a, b, c in st})


def trigger_closed_form(d2):
    """next = d2[0] if atom 1 holds the maximum, else max(d2[:k]) with k the
    first index of the maximum."""
    M = d2.max()
    k = int(np.argmax(d2))      # first occurrence
    nxt = M if k == 0 else d2[:k].max()
    return M + 2 * np.sqrt(M * nxt) + nxt


def test_trigger_closed_form(oracle):
    f = oracle.lib.oracle_maximum_approach_sq
    f.restype = ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    f.argtypes = [ctypes.c_int, dp, dp]
    rng = np.random.default_rng(0)
    for trial in range(50):
        N = int(rng.integers(2, 40))
        R0 = rng.standard_normal((N, 3))
        D = rng.standard_normal((N, 3)) * rng.uniform(0, 1, (N, 1))
        if trial % 3 == 0:
            D[0] = 10.0     # atom 1 holds the maximum
        R = R0 + D
        got = f(N, np.ascontiguousarray(R).ctypes.data_as(dp), np.ascontiguousarray(R0).ctypes.data_as(dp))
        dd = R - R0
        d2 = dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1] + dd[:, 2] * dd[:, 2]
        assert got == pytest.approx(trigger_closed_form(d2), rel=1e-15)
    # a plausible mistake -- using the true second largest -- differs here:
    d2 = np.array([0.01, 0.04, 0.03])
    assert trigger_closed_form(d2) == pytest.approx(0.04 + 2 * np.sqrt(0.04 * 0.01) + 0.01)
