"""Throughput of the EmDee force engine on one MI355X.

Main workload (large LJ, plain FP64 against FP32 pair arithmetic with FP64
accumulation): Lennard-Jones fluid of 1,048,576 atoms (fcc 64^3 cells) at density 0.8442,
T = 1.0, lj_cut with rc = 2.5 and skin 0.3, velocity Verlet dt = 0.005 driven
exactly as the reference hosts drive it (boost dt/2, displace dt, boost dt/2
through the C ABI; energies every step).  Both arithmetic forms run: plain
FP64 and FP32 pair arithmetic with FP64 accumulation; `value` is the mixed
form, reported with its force RMS error against the FP64 forces on identical
coordinates and its ratio to the plain FP64 form.

The roofline entry is the pair kernel: algorithmic bytes per launch are
N * (4 * n_half + 24 + 24) -- the int32 half Verlet list, one FP64 read of the
positions and one FP64 write of the forces per atom -- divided by the
kernel's mean duration from HIP events bound to its launches on the system's
stream (hipExtLaunchKernelGGL), taken over a second run of the same length so
that the timed run carries no profiling.

cpu_baseline times the CPU oracle (the plain-C restatement of the reference
algorithm, OpenMP over the host cores) on the same system for a bounded
number of steps.

Also reported: batched HMC of 512 LJ replicas (effective samples/s), rigid
SPC/E water of 98,304 atoms, two-layer water (exact, r-RESPA 1:4 and
short-layer step costs), and the learned-electrostatics workload on 4096
water molecules (water_hmc: HMC with the learned model as proposal force
against exact-force HMC, effective samples/s of each and their ratio, the
model's cost against the Ewald sum's, held-out force R^2, r-RESPA with the
model as fast force).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--natoms N] [--no-cpu] [--no-hmc] [--no-water]
                    [--no-water-hmc]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s vs plain FP64 at stated force error; HMC effective samples/s"
HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_pairs_lj1m_traffic.txt")


def measured_traffic(form):
    """HBM bytes per launch of the pair kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (calibrated counters), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch[{form}]"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def fcc_lattice(natoms, density):
    nc = int(np.ceil((natoms / 4) ** (1.0 / 3.0) - 1e-9))
    # the nc^3 lattice spans the box, so a truncated lattice (natoms < 4 nc^3)
    # leaves vacancies instead of overlapping its periodic images
    L = (natoms / density) ** (1.0 / 3.0)
    a = L / nc
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(np.arange(nc), np.arange(nc), np.arange(nc), indexing="ij"), -1).reshape(-1, 3)
    R = (g[:, None, :] + basis[None, :, :]).reshape(-1, 3)[:natoms] * a
    return np.ascontiguousarray(R), L


def make_system(lib, R, L, threads, rc=2.5, skin=0.3, kT=1.0, seed=86245):
    md = lib.system(threads, 1, rc, skin, R.shape[0], None, None, None)
    lib.EmDee_set_pair_model(md, 1, 1, lib.EmDee_pair_lj_cut(1.0, 1.0), 0.0)
    lib.upload(md, "box", np.array([L]))
    lib.upload(md, "coordinates", R)
    lib.random_momenta(md, kT, True, seed)
    return md


def run_steps(lib, md, nsteps, dt):
    for _ in range(nsteps):
        md.Options.Compute = True
        lib.boost(md, 1.0, 0.0, 0.5 * dt)
        lib.displace(md, 1.0, 0.0, dt)
        lib.boost(md, 1.0, 0.0, 0.5 * dt)


def gpu_run(lib, R, L, mode, steps, warmup, dt):
    md = make_system(lib, R, L, 1)
    lib.EmDeeGPU_set_precision(md, mode)
    run_steps(lib, md, warmup, dt)
    lib.EmDeeGPU_synchronize(md)
    builds0 = md.Builds
    t0 = time.perf_counter()
    run_steps(lib, md, steps, dt)
    lib.EmDeeGPU_synchronize(md)
    t = time.perf_counter() - t0
    builds = md.Builds - builds0
    # kernel times from HIP events bound to the launches, over a second run of
    # the same length (kept out of the timed run above)
    lib.EmDeeGPU_profile(md, 1)
    run_steps(lib, md, steps, dt)
    lib.EmDeeGPU_synchronize(md)
    st = lib.kernel_stats(md)
    return md, dict(seconds=t, pair_ms=st[0] / max(st[1], 1), pair_launches=int(st[1]),
                    list_ms=st[2] / max(st[3], 1), list_builds=int(st[3]), builds=builds,
                    list_entries=st[6], energy=md.Energy.Potential + md.Kinetic.Total)


def force_error(lib, md, N):
    """RMS force error of the mixed form against FP64 on identical coordinates."""
    lib.EmDeeGPU_set_precision(md, 0)
    F64 = lib.download(md, "forces", (N, 3))
    lib.EmDeeGPU_set_precision(md, 1)
    F32 = lib.download(md, "forces", (N, 3))
    return float(np.linalg.norm(F32 - F64) / np.linalg.norm(F64))


def host_cpu():
    """CPU model, physical cores and the thread count the CPU baselines use:
    one OpenMP thread per physical core, capped by OMP_NUM_THREADS when the
    environment sets it (the GPU host's share of cores)."""
    import subprocess
    model, cores, logical = "unknown", 0, os.cpu_count() or 1
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True).stdout
        cores = len({tuple(l.split(",")[1:3]) for l in out.splitlines() if l and not l.startswith("#")})
    except OSError:
        pass
    cores = cores or logical
    cap = int(os.environ.get("OMP_NUM_THREADS", cores))
    return {"model": model, "physical_cores": cores, "logical_cpus": logical, "threads": max(1, min(cores, cap)),
            "omp_places": "cores", "omp_proc_bind": "close"}


def cpu_child(kind, threads, **kw):
    """Runs a CPU baseline in a child process whose OpenMP runtime starts
    with OMP_NUM_THREADS = threads, OMP_PLACES=cores, OMP_PROC_BIND=close."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores", OMP_PROC_BIND="close")
    code = ("import json, sys; sys.path.insert(0, %r); import bench; "
            "print(json.dumps(bench.%s(**json.loads(sys.argv[1]))))" % (ROOT, kind))
    r = subprocess.run([sys.executable, "-c", code, json.dumps(dict(kw, threads=threads))], capture_output=True,
                       text=True, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline {kind} failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_lj(natoms, steps, dt, threads):
    R, L = fcc_lattice(natoms, 0.8442)
    v, t = cpu_baseline(R, L, steps, dt, threads)
    return [v, t]


def cpu_hmc(nrep, threads, seed=1):
    from emdee_amd.hmc import replica_set
    v, t = hmc_cpu_baseline(replica_set(nrep, seed=seed), nrep, threads, seed)
    return [v, t]


def cpu_baseline(R, L, steps, dt, threads):
    from emdee_amd.emdee import EmDeeLibrary
    so = os.path.join(ROOT, "oracle", "liboracle_emdee_fast.so")
    if not os.path.exists(so):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    lib = EmDeeLibrary(so)
    md = make_system(lib, R, L, threads)
    run_steps(lib, md, 1, dt)
    t0 = time.perf_counter()
    run_steps(lib, md, steps, dt)
    t = time.perf_counter() - t0
    return R.shape[0] * steps / t, t


def hmc_bench(lib, nrep, cycles, warm, seed=1):
    """Batched HMC workload (512 LJ replicas of 256-8192 atoms, densities
    0.7-0.9, kT 0.8-2.0, rc 2.5, 10-step trajectories of dt 0.005, FP32 pair
    arithmetic): graph replays of one cycle each, the potential-energy series
    of every replica (recorded on the device) for the effective sample
    count.  At most 1024 cycles."""
    from emdee_amd.hmc import BatchedHMC, integrated_autocorrelation, replica_set
    N, L, T, Rs = replica_set(nrep, seed=seed)
    h = BatchedHMC(lib, N, L, T, Rs, dt=0.005, nsteps=10, seed=seed + 6, precision=1)
    tune = h.tune()
    h.run(launches=warm, cycles=1)
    h.synchronize()
    _, acc0 = h.energies()
    # one graph launch per cycle, queued back to back; the potential energy
    # of every cycle is recorded on the device and read after the timing
    t0 = time.perf_counter()
    h.run(launches=cycles, cycles=1)
    h.synchronize()
    t = time.perf_counter() - t0
    _, acc = h.energies()
    series = h.history(cycles)
    h.close()
    if len(series) != cycles:
        raise RuntimeError("energy history shorter than the timed cycles")
    taus = np.array([integrated_autocorrelation(series[:, r]) for r in range(nrep)])
    ess = float(np.sum(cycles / taus))
    accepted = int(np.sum(acc - acc0))
    atom_steps = float(np.sum(N)) * 10 * cycles
    return dict(replicas=nrep, atoms=int(np.sum(N)), cycles=cycles, warmup_cycles=warm, seconds=t,
                atom_steps_per_s=atom_steps / t, accepted_per_s=accepted / t,
                acceptance=accepted / (cycles * nrep), effective_samples_per_s=ess / t,
                median_tau_cycles=float(np.median(taus)), launch_shape=tune), (N, L, T, Rs)


def hmc_cpu_baseline(replicas, sample, threads, seed=1):
    """Reference-API replay (one tEmDee per replica) on the oracle, for the
    first `sample` replicas and one cycle."""
    from emdee_amd.emdee import EmDeeLibrary
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import hmc_ref
    lib = EmDeeLibrary(os.path.join(ROOT, "oracle", "liboracle_emdee_fast.so"))
    N, L, T, Rs = replicas
    idx = list(range(sample))
    t0 = time.perf_counter()
    hmc_ref.replay(lib, lib, N[idx], L[idx], T[idx], [Rs[i] for i in idx], 1, seed=seed + 6, threads=threads)
    t = time.perf_counter() - t0
    return float(np.sum(N[idx])) * 10 / t, t


def water_bench(lib, nmol, steps, warm, mode, dt=2.0, seed=1):
    """Rigid SPC/E water (32,768 molecules, 98,304 atoms, L = 99.3 A):
    coul_damped_smoothed(0.2, 1 A), rc 10 A, skin 2 A, exact rotations,
    velocity Verlet through the C ABI with energies every step.  Also the
    mixed form's force RMS error against FP64 on the final coordinates."""
    from emdee_amd.water_box import KB, water_system
    md, cfg = water_system(lib, nmol, seed, coul="damped_smoothed", alpha=0.2)
    lib.EmDeeGPU_set_precision(md, mode)
    lib.random_momenta(md, KB * 298.0, True, seed)
    run_steps(lib, md, warm, dt)
    lib.EmDeeGPU_synchronize(md)
    t0 = time.perf_counter()
    run_steps(lib, md, steps, dt)
    lib.EmDeeGPU_synchronize(md)
    t = time.perf_counter() - t0
    err = force_error(lib, md, cfg["N"])
    lib.EmDeeGPU_destroy(ctypes.byref(md))
    return cfg["N"] * steps / t, 1e3 * t / steps, err


def lj_4m_bench(lib, steps, warm, dt=0.005):
    """LJ at 4,194,304 atoms (fcc lattice of 102^3 cells truncated to N,
    spanning the box), both forms, with the mixed form's force error."""
    N = 4194304
    R, L = fcc_lattice(N, 0.8442)
    out = {"natoms": N}
    for mode, name in ((0, "fp64"), (1, "mixed")):
        md, r = gpu_run(lib, R, L, mode, steps, warm, dt)
        out[name + "_atom_steps_per_s"] = N * steps / r["seconds"]
        out[name + "_pair_kernel_ms"] = r["pair_ms"]
        out[name + "_list_builds"] = r["builds"]
        if mode == 1:
            out["force_rms_error"] = force_error(lib, md, N)
        lib.EmDeeGPU_destroy(ctypes.byref(md))
    out["mixed_kernel_over_fp64_kernel"] = out["fp64_pair_kernel_ms"] / out["mixed_pair_kernel_ms"]
    out["steps"] = steps
    return out


def layers_bench(lib, nmol, outer, seed=1):
    """Rigid SPC/E water, 4096 molecules, two layers (1: coul_long + Ewald,
    2: its real-space part): milliseconds per femtosecond of trajectory for
    velocity Verlet on layer 1 (dt 1 fs), r-RESPA 1:4 (outer 4 fs, inner 1 fs
    on layer 2) and velocity Verlet on layer 2 alone."""
    from emdee_amd import mts
    from emdee_amd.surrogate import two_layer_water
    from emdee_amd.water_box import KB
    out = {}
    for name in ("exact_vv", "respa_1_4", "short_vv"):
        md, cfg = two_layer_water(lib, nmol, seed)
        lib.random_momenta(md, KB * 298.0, True, seed)
        def go(k):
            if name == "respa_1_4":
                mts.respa(lib, md, k, 4.0, 4)
            else:
                lib.switch_model_layer(md, 1 if name == "exact_vv" else 2)
                run_steps(lib, md, 4 * k, 1.0)
        go(2)
        lib.EmDeeGPU_synchronize(md)
        t0 = time.perf_counter()
        go(outer)
        lib.EmDeeGPU_synchronize(md)
        out[name + "_ms_per_fs"] = 1e3 * (time.perf_counter() - t0) / (4 * outer)
        lib.EmDeeGPU_destroy(ctypes.byref(md))
    out["respa_speedup_over_exact"] = out["exact_vv_ms_per_fs"] / out["respa_1_4_ms_per_fs"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--natoms", type=int, default=1048576)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--hmc-replicas", type=int, default=512)
    ap.add_argument("--hmc-cycles", type=int, default=400)
    ap.add_argument("--no-hmc", action="store_true")
    ap.add_argument("--no-water", action="store_true")
    ap.add_argument("--no-water-hmc", action="store_true")
    ap.add_argument("--no-4m", action="store_true")
    ap.add_argument("--steps-4m", type=int, default=100)
    ap.add_argument("--water-hmc-cycles", type=int, default=120)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the engine runs one system per GPU)", file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("EMDEE_QUIET", "1")
    import emdee_amd
    lib = emdee_amd.load()
    N, dt = args.natoms, 0.005
    R, L = fcc_lattice(N, 0.8442)

    print(f"# fp64 run: N={N} steps={args.steps} warmup={args.warmup}", file=sys.stderr, flush=True)
    md64, r64 = gpu_run(lib, R, L, 0, args.steps, args.warmup, dt)
    print(f"# fp64: {r64}", file=sys.stderr, flush=True)
    print("# mixed run", file=sys.stderr, flush=True)
    mdmx, rmx = gpu_run(lib, R, L, 1, args.steps, args.warmup, dt)
    print(f"# mixed: {rmx}", file=sys.stderr, flush=True)
    err = force_error(lib, mdmx, N)
    print(f"# force rms error (mixed vs fp64) = {err:.3e}", file=sys.stderr, flush=True)

    v64 = N * args.steps / r64["seconds"]
    vmx = N * args.steps / rmx["seconds"]
    use_mixed = err <= 1e-5
    best = rmx if use_mixed else r64
    value = vmx if use_mixed else v64
    n_half = best["list_entries"] / 2 / N
    bytes_per_launch = N * (4 * n_half + 24 + 24)
    achieved = bytes_per_launch / (best["pair_ms"] * 1e-3) / 1e9

    lj4m = None
    if not args.no_4m:
        print("# LJ 4,194,304 atoms, both forms", file=sys.stderr, flush=True)
        lj4m = lj_4m_bench(lib, args.steps_4m, 10, dt)
        print(f"# lj_4m: {lj4m}", file=sys.stderr, flush=True)

    cpu = None
    host = host_cpu()
    if not args.no_cpu:
        threads = host["threads"]
        print(f"# cpu baseline: oracle, {threads} threads, {args.cpu_steps} steps", file=sys.stderr, flush=True)
        cv, ct = cpu_child("cpu_lj", threads, natoms=N, steps=args.cpu_steps, dt=dt)
        cpu = {"value": cv, "unit": "atom-steps/s", "cores": threads, "kind": "port",
               "build": "oracle/liboracle_emdee_fast.so (-Ofast)", "host": host,
               "sample": f"{N}-atom LJ, {args.cpu_steps} steps after 1 warm-up step ({ct:.1f} s), same set-up"}

    hmc = None
    if not args.no_hmc:
        print(f"# batched HMC: {args.hmc_replicas} replicas, {args.hmc_cycles} cycles", file=sys.stderr, flush=True)
        hmc, reps = hmc_bench(lib, args.hmc_replicas, args.hmc_cycles, warm=20)
        print(f"# hmc: {hmc}", file=sys.stderr, flush=True)
        if not args.no_cpu:
            threads = host["threads"]
            hv, ht = cpu_child("cpu_hmc", threads, nrep=len(reps[0]))
            hmc["cpu_baseline"] = {"value": hv, "unit": "atom-steps/s", "cores": threads, "kind": "port",
                                   "host": host,
                                   "sample": f"reference-API replay of all {len(reps[0])} replicas, one 10-step "
                                             f"cycle each ({ht:.1f} s)"}

    water = None
    if not args.no_water:
        print("# rigid water: 32768 molecules; two-layer water: 4096 molecules", file=sys.stderr, flush=True)
        w64, m64, _ = water_bench(lib, 32768, 100, 10, 0)
        wmx, mmx, werr = water_bench(lib, 32768, 100, 10, 1)
        water = {"rigid_spce_98304": {"fp64_atom_steps_per_s": w64, "fp64_ms_per_step": m64,
                                      "mixed_atom_steps_per_s": wmx, "mixed_ms_per_step": mmx,
                                      "mixed_force_rms_error": werr,
                                      "coul": "damped_smoothed(0.2, 1.0)", "rc": 10.0, "skin": 2.0, "dt_fs": 2.0},
                 "two_layer_spce_12288": layers_bench(lib, 4096, 10)}
        print(f"# water: {water}", file=sys.stderr, flush=True)

    water_hmc = None
    if not args.no_water_hmc:
        print("# learned-electrostatics water HMC: 4096 molecules", file=sys.stderr, flush=True)
        from emdee_amd.surrogate import water_hmc_workload
        water_hmc = water_hmc_workload(lib, 4096, kcut=1.2, cycles=args.water_hmc_cycles,
                                       log=lambda *a: print("#", *a, file=sys.stderr, flush=True))

    line = {
        "metric": METRIC, "value": value, "unit": "atom-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * best["seconds"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-pair/f64-accumulate" if use_mixed else "f64", "data": "synthetic",
        "config": {"workload": "lj_fluid_fcc", "natoms": N, "density": 0.8442, "T": 1.0, "rc": 2.5,
                   "skin": 0.3, "dt": dt, "integrator": "boost/displace/boost via C ABI, energies every step",
                   "fp64_value": v64, "mixed_value": vmx, "mixed_over_fp64": vmx / v64,
                   "fp64_pair_kernel_ms": r64["pair_ms"], "mixed_pair_kernel_ms": rmx["pair_ms"],
                   "mixed_kernel_over_fp64_kernel": r64["pair_ms"] / rmx["pair_ms"],
                   "list_build_ms": best["list_ms"],
                   "force_rms_error": err, "list_builds": best["builds"], "half_list_pairs_per_atom": n_half,
                   "lj_4194304": lj4m},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("mixed" if use_mixed else "fp64"),
                     "kernel": "k_pairs_mx" if use_mixed else "k_pairs", "kernel_ms": best["pair_ms"],
                     "bytes_per_launch": bytes_per_launch,
                     "traffic_source": "profiles/k_pairs_lj1m_traffic.txt"},
        "cpu_baseline": cpu,
        "hmc": hmc,
        "water": water,
        "water_hmc": water_hmc,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
