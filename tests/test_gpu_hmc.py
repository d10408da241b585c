"""Batched HMC on the GPU (EmDeeHMC_*) against the reference-API replay of
the same runs on the oracle: one tEmDee per replica, the same momenta and
Metropolis uniforms, boost/displace/boost trajectories."""
import numpy as np
import pytest

import hmc_ref
from emdee_amd.hmc import BatchedHMC, replica_set

pytestmark = pytest.mark.gpu

CYCLES, NSTEPS, DT, SEED = 4, 10, 0.005, 77


def replicas():
    # sizes from 108 (one cell holds every atom) to 1600 (cell lists)
    return replica_set(6, seed=3, nmin=108, nmax=1600)


def test_plain_form_matches_reference_api(gpu, oracle):
    N, L, T, Rs = replicas()
    h = BatchedHMC(gpu, N, L, T, Rs, dt=DT, nsteps=NSTEPS, seed=SEED, precision=0)
    h.run(launches=1, cycles=CYCLES)
    Rg, Ug, Ag, dHg = h.state()
    h.close()
    Ro, Uo, Ao, dHo = hmc_ref.replay(oracle, oracle, N, L, T, Rs, CYCLES, dt=DT, nsteps=NSTEPS, seed=SEED)
    assert np.array_equal(Ag, Ao)
    assert np.allclose(Ug, Uo, rtol=1e-9, atol=1e-9)
    assert np.allclose(dHg, dHo[:, -1], rtol=1e-6, atol=1e-8)
    for a, b in zip(Rg, Ro):
        assert np.max(np.abs(a - b)) < 1e-8


def test_mixed_form_close_to_reference_api(gpu, oracle):
    N, L, T, Rs = replicas()
    h = BatchedHMC(gpu, N, L, T, Rs, dt=DT, nsteps=NSTEPS, seed=SEED, precision=1)
    h.run(launches=1, cycles=CYCLES)
    Rg, Ug, Ag, dHg = h.state()
    h.close()
    Ro, Uo, Ao, dHo = hmc_ref.replay(oracle, oracle, N, L, T, Rs, CYCLES, dt=DT, nsteps=NSTEPS, seed=SEED)
    print("accepted", Ag, Ao, "U rel err", np.max(np.abs(Ug - Uo) / np.abs(Uo)))
    assert np.sum(Ag != Ao) <= 1
    same = Ag == Ao
    assert np.allclose(Ug[same], Uo[same], rtol=1e-5)


def test_graph_replays_and_tuning(gpu):
    N, L, T, Rs = replicas()
    a = BatchedHMC(gpu, N, L, T, Rs, dt=DT, nsteps=NSTEPS, seed=SEED, precision=1)
    info = a.tune()
    print("autotuner", info)
    assert info["variant"] >= 0 and min(t for t in info["timings_ms"] if t >= 0) > 0
    a.run(launches=2, cycles=2)        # one graph, replayed twice
    b = BatchedHMC(gpu, N, L, T, Rs, dt=DT, nsteps=NSTEPS, seed=SEED, precision=1)
    b.run(launches=1, cycles=4)        # default launch shape, one launch
    Ra, Ua, Aa, _ = a.state()
    Rb, Ub, Ab, _ = b.state()
    a.close(); b.close()
    assert np.array_equal(Aa, Ab)
    assert np.allclose(Ua, Ub, rtol=1e-12)
    for x, y in zip(Ra, Rb):
        assert np.max(np.abs(x - y)) < 1e-10


def test_energy_history_matches_per_cycle_reads(gpu):
    """The potential energies recorded on the device after every cycle equal
    those read back after each cycle (same run, same cycles)."""
    N, L, T, Rs = replicas()
    a = BatchedHMC(gpu, N, L, T, Rs, dt=DT, nsteps=NSTEPS, seed=SEED, precision=1)
    per_cycle = []
    for _ in range(5):
        a.run(launches=1, cycles=1)
        per_cycle.append(a.energies()[0])
    hist = a.history(5)
    a.close()
    assert hist.shape == (5, len(N))
    assert np.array_equal(hist, np.array(per_cycle))
    b = BatchedHMC(gpu, N, L, T, Rs, dt=DT, nsteps=NSTEPS, seed=SEED, precision=1)
    b.run(launches=5, cycles=1)            # queued back to back, read once
    assert np.allclose(b.history(5), hist, rtol=1e-12)
    assert len(b.history(9)) == 5          # no more than the cycles run
    b.close()


def test_full_size_replica_set(gpu):
    """The 512-replica set (256 to 8192 atoms): every replica advances and
    the acceptance rate is that of a well-tuned HMC."""
    N, L, T, Rs = replica_set(512, seed=1)
    h = BatchedHMC(gpu, N, L, T, Rs, dt=DT, nsteps=NSTEPS, seed=SEED, precision=1)
    h.tune()
    h.run(launches=2, cycles=2)
    R, U, A, dH = h.state()
    h.close()
    assert np.all(np.isfinite(U)) and np.all(U / N < 0)
    assert 0.5 < A.sum() / (4 * len(N)) <= 1.0
