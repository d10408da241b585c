"""Large Lennard-Jones systems against the oracle on identical inputs.

* 1,048,576 atoms (fcc 64^3 cells at density 0.8442, the large-LJ workload's
  size), plain lj_cut, in both arithmetic forms: forces, energy, virial
  before and after ten steps with list rebuilds, kinetic energy and rebuild
  count, and the Verlet-list pair set entry for entry (systems this large
  take the one-lane list build).  The oracle runs with all host cores (up to
  16 threads).
* 131,072 atoms with shifted-force LJ, which no dedicated LJ kernel serves:
  the pair loop takes the generic kernels with one thread per atom (short
  lists at this size), in both forms.
"""
import ctypes
import os

import numpy as np
import pytest

from test_gpu_mixed import FORCE_RMS, errors, fcc_box

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def make(lib, R, L, model, gpu_mixed, threads=1, exclusions=()):
    N = len(R)
    md = lib.system(threads, 1, 2.5, 0.3, N, None, None, None)
    lib.EmDee_set_pair_model(md, 1, 1, model(lib), 0.0)
    for i, j in exclusions:
        lib.EmDee_ignore_pair(md, i, j)
    if gpu_mixed:
        lib.EmDeeGPU_set_precision(md, 1)
    lib.upload(md, "box", np.array([L]))
    lib.upload(md, "coordinates", R)
    lib.random_momenta(md, 1.0, True, 5)
    return md


def pair_keys(pairs, N):
    """Sorted int64 keys (i - 1) * N + (j - 1) of 1-based pairs i < j."""
    p = np.asarray(pairs, dtype=np.int64)
    return np.sort((p[:, 0] - 1) * N + (p[:, 1] - 1))


def oracle_pair_array(oracle, md):
    f = oracle.lib.oracle_neighbor_pairs
    f.restype = ctypes.c_int
    f.argtypes = [type(md), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    n = f(md, 0, None)
    out = np.zeros((max(n, 1), 2), dtype=np.int32)
    f(md, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return out[:n]


def compare(gpu, oracle, nc, model, mixed, steps, label, jitter, pairs=False):
    R, L = fcc_box(nc, 0.8442, jitter, seed=23)
    N = len(R)
    mg = make(gpu, R, L, model, mixed)
    mo = make(oracle, R, L, model, False, THREADS)
    tol = FORCE_RMS if mixed else 1e-9
    for stage in range(2):
        if stage:
            for _ in range(steps):
                for lib, md in ((gpu, mg), (oracle, mo)):
                    lib.boost(md, 1.0, 0.0, 0.0025)
                    lib.displace(md, 1.0, 0.0, 0.005)
                    lib.boost(md, 1.0, 0.0, 0.0025)
        gpu.compute_forces(mg)
        oracle.compute_forces(mo)
        ferr, eerr, werr = errors(gpu, oracle, mg, mo, N)
        dW = abs(mg.Virial - mo.Virial) / N
        kerr = abs(mg.Kinetic.Total - mo.Kinetic.Total) / mo.Kinetic.Total
        print(f"{label} N={N} mixed={mixed} stage {stage}: force rms {ferr:.2e} energy {eerr:.2e} "
              f"virial per atom {dW:.1e} kinetic {kerr:.1e} builds {mg.Builds}/{mo.Builds}", flush=True)
        assert ferr < tol and eerr < 1e-6 and dW < 2e-5 and kerr < 1e-6
        assert mg.Builds == mo.Builds
        if pairs:
            kg, ko = pair_keys(gpu.neighbor_pairs(mg), N), pair_keys(oracle_pair_array(oracle, mo), N)
            print(f"{label} stage {stage}: {len(kg)} pairs on the GPU, {len(ko)} in the oracle", flush=True)
            assert len(kg) == len(ko) and np.array_equal(kg, ko)
    gpu.EmDeeGPU_destroy(mg)


@pytest.mark.parametrize("mixed", [False, True])
def test_million_atom_lj(gpu, oracle, mixed):
    # a Gaussian jitter of 0.08 sigma would bring some of the 6.3 million
    # nearest-neighbour pairs within 0.6 sigma (forces ~2e4) at this size
    compare(gpu, oracle, 64, lambda lib: lib.EmDee_pair_lj_cut(1.0, 1.0), mixed, 10, "lj_cut", 0.02, pairs=True)


@pytest.mark.parametrize("mixed", [False, True])
def test_generic_one_thread_kernels(gpu, oracle, mixed):
    compare(gpu, oracle, 32, lambda lib: lib.EmDee_shifted_force(lib.EmDee_pair_lj_cut(1.0, 1.0)), mixed, 20,
            "lj_sf", 0.08)


def cluster_box(nc, seed):
    """An fcc box with its first 392 atoms moved onto a 7 x 7 x 8 grid of
    spacing 0.3 sigma at the centre: those atoms hold about 470 list entries
    each, about four times the list capacity the build starts from."""
    R, L = fcc_box(nc, 0.8442, 0.02, seed=seed)
    g = np.arange(7) * 0.3
    cl = np.stack(np.meshgrid(g, g, np.arange(8) * 0.3, indexing="ij"), -1).reshape(-1, 3) + 0.5 * L + 0.05
    R = np.array(R, copy=True)
    R[: len(cl)] = cl
    return R, L


@pytest.mark.parametrize("nc,mixed,excl", [(10, False, False), (10, True, False), (10, False, True), (10, True, True),
                                           (64, False, False), (64, True, False)])
def test_list_capacity_overflow(gpu, oracle, nc, mixed, excl):
    # 4,000 atoms take the 16-lane build, 1,048,576 the one-lane segment
    # build: both must drop the entries past the capacity, report the
    # longest list and be rebuilt with room for it.  With exclusions (two
    # pairs inside the cluster) the 16-lane build takes its filtered form.
    R, L = cluster_box(nc, 29)
    N = len(R)
    model = lambda lib: lib.EmDee_pair_lj_cut(1.0, 1.0)
    ex = [(1, 2), (5, 60)] if excl else []
    mg = make(gpu, R, L, model, mixed, exclusions=ex)
    mo = make(oracle, R, L, model, False, THREADS, exclusions=ex)
    gpu.compute_forces(mg)
    oracle.compute_forces(mo)
    kg, ko = pair_keys(gpu.neighbor_pairs(mg), N), pair_keys(oracle_pair_array(oracle, mo), N)
    print(f"cluster N={N} mixed={mixed} excl={excl}: {len(kg)} pairs on the GPU, {len(ko)} in the oracle", flush=True)
    assert len(kg) == len(ko) and np.array_equal(kg, ko)
    ferr, eerr, _ = errors(gpu, oracle, mg, mo, N)
    print(f"cluster N={N} mixed={mixed} excl={excl}: force rms {ferr:.2e} energy {eerr:.2e}", flush=True)
    # The cluster's 0.3 sigma pairs dominate the energy.  The mixed form's
    # positions are 32-bit fixed point (L / 2^32 = 2.5e-8 sigma at 1M atoms),
    # a relative error of 8e-8 at 0.3 sigma, which r^-12 raises about twelve
    # times: the energy bound is 1e-5 there (1e-6 for the FP64 form).
    assert ferr < (FORCE_RMS if mixed else 1e-9) and eerr < (1e-5 if mixed else 1e-6)
    gpu.EmDeeGPU_destroy(mg)
