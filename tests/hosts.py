"""Host programs for the EmDee C ABI, written after the reference's test hosts.

`read_configuration` follows test/common/mConfig.f90:41-66 and `run_md`
follows test/common/contained.f90:56-89 (boost(dt/2) / displace(dt) /
boost(dt/2), energies every `nprop` steps).  `lj_sample_system` reproduces the
set-up of test/test_pair_lj_*.f90:36-44.  Each helper takes an EmDeeLibrary,
so the same host drives the HIP product library or the CPU oracle.
"""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_configuration(path):
    """Parse a NIST .mol fixture (mConfig.f90:41-66)."""
    with open(path) as f:
        lines = f.read().splitlines()
    it = iter(lines)
    next(it); next(it)                        # description
    next(it); ntypes = int(next(it).split()[0])
    next(it)
    mass = np.zeros(ntypes); eps = np.zeros(ntypes); sig = np.zeros(ntypes)
    for _ in range(ntypes):
        p = next(it).split()
        k = int(p[0]) - 1
        mass[k], eps[k], sig[k] = float(p[1]), float(p[2]), float(p[3])
    next(it); L = float(next(it).split()[0])
    next(it); N = int(next(it).split()[0])
    next(it)
    R = np.zeros((N, 3)); types = np.zeros(N, np.int32); mol = np.zeros(N, np.int32); Q = np.zeros(N)
    for _ in range(N):
        p = next(it).split()
        k = int(p[0]) - 1
        mol[k] = int(p[1]); types[k] = int(p[2]); Q[k] = float(p[3])
        R[k] = [float(x) for x in p[4:7]]
    return dict(ntypes=ntypes, mass=mass, epsilon=eps, sigma=sig, L=L, N=N, R=R,
                types=types, molecule=mol, Q=Q)


def load_pins():
    with open(os.path.join(GOLDEN, "lj_pins.json")) as f:
        return json.load(f)


def lj_model(lib, name, eps, sig, rc):
    base = lib.EmDee_pair_lj_cut(eps, sig)
    if name == "lj_cut":
        return base
    if name == "lj_sf":
        return lib.EmDee_shifted_force(base)
    if name == "lj_smoothed":
        return lib.EmDee_square_smoothed(base, rc - 1.0)
    if name == "lj_shifted_smoothed":
        return lib.EmDee_shifted_square_smoothed(base, rc - 1.0)
    raise ValueError(name)


def lj_sample_system(lib, name, threads=1):
    """Set-up of test/test_pair_lj_*.f90:36-44 on NIST_lj_sample.mol."""
    pins = load_pins()["input"]
    cfg = read_configuration(os.path.join(GOLDEN, "NIST_lj_sample.mol"))
    rc, skin = pins["rc"], pins["skin"]
    md = lib.system(threads, 1, rc, skin, cfg["N"], cfg["types"], cfg["mass"], None)
    for i in range(cfg["ntypes"]):
        model = lj_model(lib, name, cfg["epsilon"][i], cfg["sigma"][i], rc)
        lib.EmDee_set_pair_model(md, i + 1, i + 1, model, pins["kCoul"])
    lib.upload(md, "charges", cfg["Q"])
    lib.upload(md, "box", np.array([cfg["L"]]))
    lib.upload(md, "coordinates", cfg["R"])
    lib.random_momenta(md, pins["kB"] * pins["temp"], True, pins["seed"])
    return md, cfg


def run_md(lib, md, nsteps, dt, nprop):
    """contained.f90:56-89: returns [U, W, U+K] after nsteps."""
    for step in range(1, nsteps + 1):
        md.Options.Compute = (step % nprop) == 0
        lib.boost(md, 1.0, 0.0, 0.5 * dt)
        lib.displace(md, 1.0, 0.0, dt)
        lib.boost(md, 1.0, 0.0, 0.5 * dt)
    return np.array([md.Energy.Potential, md.Virial, md.Energy.Potential + md.Kinetic.Total])
