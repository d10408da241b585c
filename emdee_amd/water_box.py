"""Synthetic rigid 3-site water boxes (SPC/E geometry and parameters) for
the water workloads: molecules on a simple-cubic lattice of centres with
random orientations from a seeded generator, at 1.0 g/cm^3.

Units are those of the reference's SPC/E sample (test/spce_sample.inp):
Da, Angstrom, fs, charges in e; energies in Da A^2 / fs^2.
"""
import numpy as np

KB = 8.3144621455e-7           # Da A^2 / (fs^2 K)
KCOUL = 0.13893545755135628    # Da A^3 / (fs e)^2
KJMOL = 1.0 / (4.184 * 2390.057364)   # kJ/mol in Da A^2/fs^2 (via kcal/mol)
MASS = np.array([15.9994, 1.0080])
CHARGE = np.array([-0.8476, 0.4238])
EPS_OO = 0.6502 * KJMOL
SIG_OO = 3.166
OH, HOH = 1.0, np.deg2rad(109.47)
AVOGADRO = 6.02214076e23


def box_length(nmol, density=1.0):
    m = nmol * (MASS[0] + 2 * MASS[1]) / AVOGADRO          # g
    return (m / density) ** (1.0 / 3.0) * 1e8               # A


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a*a + b*b - c*c - d*d, 2*(b*c - a*d), 2*(b*d + a*c)],
                     [2*(b*c + a*d), a*a - b*b + c*c - d*d, 2*(c*d - a*b)],
                     [2*(b*d - a*c), 2*(c*d + a*b), a*a - b*b - c*c + d*d]])


def water_box(nmol, seed, density=1.0):
    """Returns (R [3 nmol, 3], types [3 nmol] (1 = O, 2 = H), molecule ids
    [3 nmol], box length)."""
    rng = np.random.default_rng(seed)
    L = box_length(nmol, density)
    m = int(np.ceil(nmol ** (1.0 / 3.0) - 1e-9))
    g = np.stack(np.meshgrid(*(np.arange(m),) * 3, indexing="ij"), -1).reshape(-1, 3)[:nmol]
    centres = (g + 0.5) * (L / m)
    # body-frame geometry: O at origin, H in the xy plane
    hx, hy = OH * np.cos(HOH / 2), OH * np.sin(HOH / 2)
    body = np.array([[0.0, 0.0, 0.0], [hx, hy, 0.0], [hx, -hy, 0.0]])
    com = (MASS[0] * body[0] + MASS[1] * (body[1] + body[2])) / (MASS[0] + 2 * MASS[1])
    body -= com
    R = np.empty((3 * nmol, 3))
    for k in range(nmol):
        R[3 * k: 3 * k + 3] = centres[k] + body @ random_rotation(rng).T
    types = np.tile(np.array([1, 2, 2], dtype=np.int32), nmol)
    mol = np.repeat(np.arange(1, nmol + 1, dtype=np.int32), 3)
    return R, types, mol, L


def water_system(lib, nmol, seed, rc=10.0, skin=2.0, coul="long", accuracy=np.exp(-9.0) / 9.0, alpha=0.3,
                 layers=1, mode=0):
    """An EmDee system of `nmol` rigid SPC/E molecules.  coul: "long"
    (coul_long + kspace_ewald(accuracy); alpha = 0.3/A at rc = 10 A for the
    default accuracy), "damped" (coul_damped(alpha): the real-space part of
    coul_long alone), "damped_smoothed" (coul_damped_smoothed(alpha, 1 A))."""
    R, types, mol, L = water_box(nmol, seed)
    md = lib.system(1, layers, rc, skin, len(types), types, MASS, mol)
    lib.EmDee_set_pair_model(md, 1, 1, lib.EmDee_pair_lj_cut(EPS_OO, SIG_OO), KCOUL)
    lib.EmDee_set_pair_model(md, 2, 2, lib.EmDee_pair_none(), KCOUL)
    if coul == "long":
        lib.EmDee_set_coul_model(md, lib.EmDee_coul_long())
        lib.EmDee_set_kspace_model(md, lib.EmDee_kspace_ewald(accuracy))
    elif coul == "damped":
        lib.EmDee_set_coul_model(md, lib.EmDee_coul_damped(alpha))
    else:
        lib.EmDee_set_coul_model(md, lib.EmDee_coul_damped_smoothed(alpha, 1.0))
    md.Options.RotationMode = mode
    lib.upload(md, "charges", CHARGE[types - 1])
    lib.upload(md, "coordinates", R)
    lib.upload(md, "box", np.array([L]))
    return md, dict(R=R, types=types, mol=mol, L=L, N=len(types))


if __name__ == "__main__":
    # Profiling workload: python -m emdee_amd.water_box [molecules] [steps] [mixed]
    # rigid SPC/E, coul_damped_smoothed(0.2, 1 A), rc 10 A, skin 2 A, dt 2 fs
    import os
    import sys
    import time

    from . import load
    nmol = int(sys.argv[1]) if len(sys.argv) > 1 else 32768
    steps = int(sys.argv[2]) if len(sys.argv) > 2 else 20
    mixed = int(sys.argv[3]) if len(sys.argv) > 3 else 0
    os.environ.setdefault("EMDEE_QUIET", "1")
    lib = load()
    md, cfg = water_system(lib, nmol, 1, coul="damped_smoothed", alpha=0.2)
    lib.EmDeeGPU_set_precision(md, mixed)
    lib.random_momenta(md, KB * 298.0, True, 1)
    t0 = time.perf_counter()
    for _ in range(steps):
        md.Options.Compute = True
        lib.boost(md, 1.0, 0.0, 1.0)
        lib.displace(md, 1.0, 0.0, 2.0)
        lib.boost(md, 1.0, 0.0, 1.0)
    lib.EmDeeGPU_synchronize(md)
    t = time.perf_counter() - t0
    print(f"{cfg['N']} atoms, {steps} steps: {1e3 * t / steps:.3f} ms/step, builds {md.Builds}", flush=True)
