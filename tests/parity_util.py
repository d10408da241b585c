"""Shared helpers of the parity tests: synthetic systems, brute-force pair
sets and library-agnostic set-up, so that one host function drives both the
HIP library and the CPU oracle on identical inputs."""
import ctypes

import numpy as np


def fcc(natoms, density, jitter=0.0, seed=0):
    nc = int(np.ceil((natoms / 4) ** (1.0 / 3.0) - 1e-9))
    a = (4.0 / density) ** (1.0 / 3.0)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(np.arange(nc), np.arange(nc), np.arange(nc), indexing="ij"), -1).reshape(-1, 3)
    R = (g[:, None, :] + basis[None, :, :]).reshape(-1, 3)[:natoms] * a
    L = (natoms / density) ** (1.0 / 3.0)
    if jitter:
        R = R + jitter * np.random.default_rng(seed).standard_normal(R.shape)
    return np.ascontiguousarray(R), L


def random_box(natoms, L, seed, min_dist=0.8):
    """Random positions with a minimum separation (rejection sampling)."""
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < natoms:
        p = rng.uniform(-0.5 * L, 1.5 * L, 3)   # deliberately outside [0, L)
        ok = True
        for q in pts[-200:]:
            d = (p - q) / L
            d -= np.round(d)
            if np.sum((d * L) ** 2) < min_dist ** 2:
                ok = False
                break
        if ok:
            pts.append(p)
    return np.array(pts)


def brute_pairs(R, L, cutoff, exclude=None, types=None, interact=None):
    """All pairs i < j (1-based) with the reference's FP64 test
    r^2 < cutoff^2/L^2 on R/L with anint minimum image
    (neighbor_lists.f90:244-245).  numpy evaluates r^2 as dx*dx + dy*dy + dz*dz
    without contraction; np.round differs from anint only at exact ties,
    where |d - round(d)| is the same."""
    Rs = R / L
    c2 = (cutoff * cutoff) / (L * L)
    N = len(R)
    out = []
    for i in range(N - 1):
        d = Rs[i] - Rs[i + 1:]
        d = d - np.round(d)
        r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        js = np.nonzero(r2 < c2)[0] + i + 1
        for j in js:
            if exclude and (i + 1, j + 1) in exclude:
                continue
            if types is not None and interact is not None and not interact[types[i] - 1][types[j] - 1]:
                continue
            out.append((i + 1, j + 1))
    return set(out)


def oracle_pairs(lib, md):
    f = lib.lib.oracle_neighbor_pairs
    f.restype = ctypes.c_int
    f.argtypes = [type(md), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    n = f(md, 0, None)
    out = np.zeros((max(n, 1), 2), dtype=np.int32)
    f(md, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return set(map(tuple, out[:n].tolist()))


def gpu_pairs(lib, md):
    return set(map(tuple, lib.neighbor_pairs(md).tolist()))


def build(lib, R, L, rc, skin, models, types=None, masses=None, charges=None, coul=None,
          kcoul=None, exclusions=(), threads=1, compute=True, cross=None):
    """Create a system, set models and upload box/coordinates.

    models: list of pair-model handles constructor callables per type
    (lib -> handle); cross: {(i, j): ctor} explicit cross pairs."""
    N = len(R)
    md = lib.system(threads, 1, rc, skin, N, types, masses, None)
    md.Options.Compute = compute
    for t, ctor in enumerate(models):
        lib.EmDee_set_pair_model(md, t + 1, t + 1, ctor(lib), 0.0 if kcoul is None else kcoul)
    for (i, j), ctor in (cross or {}).items():
        lib.EmDee_set_pair_model(md, i, j, ctor(lib), 0.0 if kcoul is None else kcoul)
    if coul is not None:
        lib.EmDee_set_coul_model(md, coul(lib))
    for i, j in exclusions:
        lib.EmDee_ignore_pair(md, i, j)
    if charges is not None:
        lib.upload(md, "charges", charges)
    lib.upload(md, "box", np.array([L]))
    lib.upload(md, "coordinates", R)
    return md
