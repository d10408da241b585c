"""EmDeeGPU_rdf_noalloc, the allocation-free form of EmDee_rdf: bit-identical
results to EmDee_rdf through the C ABI on two calls in a row (the second one
smaller, so that stale counts of the first would show), and a clear error for
a request beyond the static scratch array."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import hosts
import water

pytestmark = pytest.mark.gpu

IP = ctypes.POINTER(ctypes.c_int)
DP = ctypes.POINTER(ctypes.c_double)


def rdf_with(fn, md, bins, rc, pairs):
    it = np.array([p[0] for p in pairs], np.int32)
    jt = np.array([p[1] for p in pairs], np.int32)
    g = np.full(bins * len(pairs), np.nan)
    fn(md, bins, rc, len(pairs), it.ctypes.data_as(IP), jt.ctypes.data_as(IP), g.ctypes.data_as(DP))
    return g.reshape(len(pairs), bins)


def test_noalloc_rdf_matches_rdf(gpu):
    md, cfg = water.spce_fixture_system(gpu, rc=9.0, skin=1.0)
    water.run(gpu, md, 5, 2.0)
    calls = [(120, 9.0, [(1, 1), (1, 2), (2, 2)]),     # first call: larger
             (40, 6.0, [(2, 1)])]                       # second: fewer bins, one pair, reversed types
    for bins, rc, pairs in calls:
        print(f"rdf_noalloc bins={bins} rc={rc} pairs={pairs}", flush=True)
        ref = rdf_with(gpu.EmDee_rdf, md, bins, rc, pairs)
        got = rdf_with(gpu.EmDeeGPU_rdf_noalloc, md, bins, rc, pairs)
        assert np.array_equal(got, ref)
        assert np.sum(ref) > 0
    # and a larger call again after the smaller one
    bins, rc, pairs = calls[0]
    assert np.array_equal(rdf_with(gpu.EmDeeGPU_rdf_noalloc, md, bins, rc, pairs),
                          rdf_with(gpu.EmDee_rdf, md, bins, rc, pairs))


def test_noalloc_rdf_lj_after_dynamics(gpu):
    md, cfg = hosts.lj_sample_system(gpu, "lj_cut")
    hosts.run_md(gpu, md, 30, 0.005, 10)
    for bins, rc in ((300, 3.0), (7, 2.0)):
        print(f"rdf_noalloc lj bins={bins} rc={rc}", flush=True)
        assert np.array_equal(rdf_with(gpu.EmDeeGPU_rdf_noalloc, md, bins, rc, [(1, 1)]),
                              rdf_with(gpu.EmDee_rdf, md, bins, rc, [(1, 1)]))


def test_noalloc_rdf_rejects_oversize(gpu):
    """bins * type-pair slots above the static array's 65536 counts ends with
    the reference's error format and exit status 1 (no kernel is launched)."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = (
        "import sys, ctypes, numpy as np; sys.path[:0] = [{root!r}]\n"
        "import emdee_amd\n"
        "lib = emdee_amd.load()\n"
        "md = lib.system(1, 1, 2.5, 0.3, 2, np.array([1, 2], np.int32), None, None)\n"
        "it = np.array([1, 1, 2], np.int32); jt = np.array([1, 2, 2], np.int32)\n"
        "g = np.zeros(3 * 30000)\n"
        "ip = ctypes.POINTER(ctypes.c_int)\n"
        "lib.EmDeeGPU_rdf_noalloc(md, 30000, 2.5, 3, it.ctypes.data_as(ip), jt.ctypes.data_as(ip),\n"
        "                         g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))\n"
        "print('not reached')\n"
    ).format(root=root)
    env = dict(os.environ, EMDEE_QUIET="1")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 1, r.stderr
    assert "Error in radial distribution calculation: more bins than the allocation-free form holds" in r.stderr
    assert "not reached" not in r.stdout
