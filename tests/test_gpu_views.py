"""EmDee_memory_address on the GPU against the oracle (EmDeeCode.f90:212-235):
the returned arrays follow the engine's state after every call, and values
written through them reach the engine at the next call, as writes through
the reference's pointers do (nothing is marked stale)."""
import ctypes

import numpy as np
import pytest

import parity_util as pu

pytestmark = pytest.mark.gpu


def view(lib, md, name, N):
    p = lib.EmDee_memory_address(md, name.encode())
    return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_double)), shape=(N, 3))


def steps(lib, md, n, dt=0.005):
    for _ in range(n):
        lib.boost(md, 1.0, 0.0, 0.5 * dt)
        lib.displace(md, 1.0, 0.0, dt)
        lib.boost(md, 1.0, 0.0, 0.5 * dt)


def test_views_follow_and_write_back(gpu, oracle):
    R, L = pu.fcc(500, 0.8442, jitter=0.05, seed=9)
    N = len(R)
    runs = []
    for lib in (gpu, oracle):
        md = pu.build(lib, R, L, 2.5, 0.3, [lambda l: l.EmDee_pair_lj_cut(1.0, 1.0)])
        md.Options.Compute = True
        lib.random_momenta(md, 1.0, True, 3)
        vr, vp = view(lib, md, "coordinates", N), view(lib, md, "momenta", N)
        steps(lib, md, 10)
        assert np.array_equal(vr, lib.download(md, "coordinates", (N, 3)))
        assert np.array_equal(vp, lib.download(md, "momenta", (N, 3)))
        vr[7, 0] += 0.05               # written through the views
        vp[11, 1] = -vp[11, 1]
        lib.compute_forces(md)
        e = (md.Energy.Potential, md.Virial)
        steps(lib, md, 10)
        runs.append((e, md.Energy.Potential + md.Kinetic.Total, np.array(vr), md.Builds))
    (eg, hg, rg, bg), (eo, ho, ro, bo) = runs
    assert eg == pytest.approx(eo, rel=1e-12)
    assert hg == pytest.approx(ho, rel=1e-10)
    assert np.max(np.abs(rg - ro)) < 1e-10
    assert bg == bo
