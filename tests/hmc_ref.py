"""Reference-API replay of the batched HMC runs: one tEmDee per replica,
driven through EmDee_upload("momenta"), EmDee_boost / EmDee_displace and a
Metropolis test, with the momenta and uniforms of the oracle's restatement of
the counter-based generator."""
import ctypes

import numpy as np

_dp = ctypes.POINTER(ctypes.c_double)


def _fn(oracle, name, res, args):
    f = getattr(oracle.lib, name)
    f.restype = res
    f.argtypes = args
    return f


def momenta(oracle, seed, replica, cycle, n, kT):
    f = _fn(oracle, "oracle_hmc_momenta", None,
            [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp])
    P = np.zeros((n, 3))
    f(seed, replica, cycle, n, kT, P.ctypes.data_as(_dp))
    return P


def accept_uniform(oracle, seed, replica, cycle):
    return _fn(oracle, "oracle_hmc_accept_uniform", ctypes.c_double,
               [ctypes.c_uint64, ctypes.c_int, ctypes.c_int])(seed, replica, cycle)


def replay(lib, oracle, N, L, T, Rs, cycles, rc=2.5, dt=0.005, nsteps=10, seed=2024, skin=0.3, threads=1):
    """Runs `cycles` HMC cycles of every replica on `lib` (any library with
    the EmDee ABI); returns final coordinates, potential energies, accepted
    counts and the energy changes of every cycle ([replica][cycle])."""
    outR, outU, outA, outD = [], [], [], []
    for r, (n, box, kT, R0) in enumerate(zip(N, L, T, Rs)):
        md = lib.system(threads, 1, rc, skin, int(n), None, None, None)
        lib.EmDee_set_pair_model(md, 1, 1, lib.EmDee_pair_lj_cut(1.0, 1.0), 0.0)
        lib.upload(md, "box", np.array([box]))
        lib.upload(md, "coordinates", R0)
        acc, dHs = 0, []
        for c in range(cycles):
            Rsave = lib.download(md, "coordinates", (int(n), 3))
            lib.upload(md, "momenta", momenta(oracle, seed, r, c, int(n), kT))
            H0 = md.Energy.Potential + md.Kinetic.Total
            for _ in range(nsteps):
                lib.boost(md, 1.0, 0.0, 0.5 * dt)
                lib.displace(md, 1.0, 0.0, dt)
                lib.boost(md, 1.0, 0.0, 0.5 * dt)
            dH = md.Energy.Potential + md.Kinetic.Total - H0
            dHs.append(dH)
            if dH <= 0.0 or accept_uniform(oracle, seed, r, c) < np.exp(-dH / kT):
                acc += 1
            else:
                lib.upload(md, "coordinates", Rsave)
                lib.compute_forces(md)   # the upload leaves energies stale
        outR.append(lib.download(md, "coordinates", (int(n), 3)))
        outU.append(md.Energy.Potential)
        outA.append(acc)
        outD.append(dHs)
    return outR, np.array(outU), np.array(outA), np.array(outD)
