"""Batched hybrid Monte Carlo of independent Lennard-Jones replicas on the
GPU (the EmDeeHMC_* entry points of include/emdee_gpu.h).

A reference user would run each replica as its own tEmDee: upload momenta,
boost(dt/2) / displace(dt) / boost(dt/2) for nsteps, then accept or reject
on the total energy (EmDeeCode.f90:805-925, 1024-1103).  `BatchedHMC` runs
all replicas at once in one persistent, graph-captured launch.
"""
import ctypes

import numpy as np

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def replica_set(nrep, seed, nmin=256, nmax=8192, rho=(0.7, 0.9), kT=(0.8, 2.0)):
    """Sizes uniform in [nmin, nmax], densities and temperatures uniform in
    the given ranges (numpy Generator seeded with `seed`); each replica
    starts on N sites drawn at random from an fcc lattice filling its box,
    slightly jittered."""
    rng = np.random.default_rng(seed)
    N = rng.integers(nmin, nmax + 1, size=nrep)
    dens = rng.uniform(*rho, size=nrep)
    T = rng.uniform(*kT, size=nrep)
    L = (N / dens) ** (1.0 / 3.0)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    Rs = []
    for n, box in zip(N, L):
        m = int(np.ceil((n / 4) ** (1.0 / 3.0) - 1e-9))
        g = np.stack(np.meshgrid(*(np.arange(m),) * 3, indexing="ij"), -1).reshape(-1, 3)
        sites = (g[:, None, :] + basis[None, :, :]).reshape(-1, 3)
        sites = sites[np.sort(rng.choice(len(sites), size=n, replace=False))]
        R = (sites + 0.25) * (box / m) + rng.uniform(-0.02, 0.02, size=(n, 3))
        Rs.append(R)
    return N.astype(np.int32), L, T, Rs


class BatchedHMC:
    def __init__(self, lib, natoms, box, kT, coords, rc=2.5, dt=0.005, nsteps=10, seed=2024, precision=1):
        self.lib = lib.lib if hasattr(lib, "lib") else lib
        self.N = np.ascontiguousarray(natoms, dtype=np.int32)
        self.L = np.ascontiguousarray(box, dtype=np.float64)
        self.kT = np.ascontiguousarray(kT, dtype=np.float64)
        R = np.ascontiguousarray(np.concatenate([np.asarray(r, dtype=np.float64).reshape(-1, 3) for r in coords]))
        self.offsets = np.concatenate([[0], np.cumsum(self.N)])
        self.seed = seed
        self.h = self.lib.EmDeeHMC_create(len(self.N), self.N.ctypes.data_as(_ip), self.L.ctypes.data_as(_dp),
                                          self.kT.ctypes.data_as(_dp), R.ctypes.data_as(_dp), rc, dt, nsteps,
                                          seed, precision)

    def tune(self):
        best = self.lib.EmDeeHMC_tune(self.h)
        bs, div = ctypes.c_int(), ctypes.c_int()
        t = np.zeros(6)
        self.lib.EmDeeHMC_variant(self.h, ctypes.byref(bs), ctypes.byref(div), t.ctypes.data_as(_dp))
        return dict(variant=best, block=bs.value, cell_division=div.value, timings_ms=t.tolist())

    def run(self, launches=1, cycles=1):
        self.lib.EmDeeHMC_run(self.h, launches, cycles)

    def synchronize(self):
        self.lib.EmDeeHMC_synchronize(self.h)

    def state(self):
        n = int(self.offsets[-1])
        R = np.zeros((n, 3))
        U = np.zeros(len(self.N))
        acc = np.zeros(len(self.N), dtype=np.int64)
        dH = np.zeros(len(self.N))
        self.lib.EmDeeHMC_get(self.h, R.ctypes.data_as(_dp), U.ctypes.data_as(_dp),
                              acc.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), dH.ctypes.data_as(_dp))
        return [R[a:b] for a, b in zip(self.offsets[:-1], self.offsets[1:])], U, acc, dH

    def energies(self):
        """Potential energies and accepted counts only (no coordinate copy)."""
        U = np.zeros(len(self.N))
        acc = np.zeros(len(self.N), dtype=np.int64)
        self.lib.EmDeeHMC_get(self.h, None, U.ctypes.data_as(_dp),
                              acc.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), None)
        return U, acc

    def history(self, cycles):
        """Potential energies after each of the last `cycles` cycles, recorded
        on the device: array (cycles, replicas), oldest first."""
        U = np.zeros((cycles, len(self.N)))
        n = self.lib.EmDeeHMC_history(self.h, cycles, U.ctypes.data_as(_dp))
        return U[:n]

    def close(self):
        if self.h:
            self.lib.EmDeeHMC_destroy(self.h)
            self.h = None


def integrated_autocorrelation(x):
    """Integrated autocorrelation time of a series by Geyer's initial
    positive sequence estimator (Geyer, Stat. Sci. 7, 1992)."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    x = x - x.mean()
    var = np.dot(x, x) / n
    if var == 0.0 or n < 4:
        return 1.0
    acf = np.array([np.dot(x[: n - k], x[k:]) / (n * var) for k in range(n // 2)])
    tau = -1.0
    for m in range(0, len(acf) - 1, 2):
        g = acf[m] + acf[m + 1]
        if g <= 0.0:
            break
        tau += 2.0 * g
    return max(tau, 1.0)


if __name__ == "__main__":
    # Profiling workload: python -m emdee_amd.hmc [replicas] [cycles]
    import sys
    import time

    from . import load
    nrep = int(sys.argv[1]) if len(sys.argv) > 1 else 512
    cycles = int(sys.argv[2]) if len(sys.argv) > 2 else 4
    lib = load()
    N, L, T, Rs = replica_set(nrep, seed=1)
    h = BatchedHMC(lib, N, L, T, Rs, seed=7, precision=1)
    print("launch shape", h.tune(), flush=True)
    t0 = time.perf_counter()
    h.run(launches=cycles, cycles=1)
    h.synchronize()
    t = time.perf_counter() - t0
    U, acc = h.energies()
    print(f"{nrep} replicas, {int(N.sum())} atoms, {cycles} cycles in {t:.3f} s; accepted {int(acc.sum())}", flush=True)
    h.close()
