"""Parity of the HIP engine with the CPU oracle on identical inputs.

* Neighbor lists (neighbor_lists.f90:199-300): the unordered pair set of the
  GPU list equals the oracle's and a brute-force all-pairs set, bit-exactly.
* Forces, energies and virials (compute.f90 + apply_modifier.f90 + model
  files): every pair model, modifier, mixing rule and Coulomb model, in both
  the energy and the virial-only mode.  FP64 tolerance: relative 1e-12
  (summation order differs: the oracle follows the reference cell order, the
  GPU sums per atom over a full list).
* Mixed form: FP32 pair arithmetic with FP64 accumulation, force RMS error
  below 1e-5 relative.
"""
import numpy as np
import pytest

import parity_util as pu

pytestmark = pytest.mark.gpu

FTOL = 1e-12


def lj(e=1.0, s=1.0):
    return lambda lib: lib.EmDee_pair_lj_cut(e, s)


def compare_state(gpu, ora, mg, mo, N, ftol=FTOL):
    Fg = gpu.download(mg, "forces", (N, 3))
    Fo = ora.download(mo, "forces", (N, 3))
    scale = np.max(np.abs(Fo))
    assert np.max(np.abs(Fg - Fo)) <= ftol * scale, np.max(np.abs(Fg - Fo)) / scale
    assert abs(mg.Virial - mo.Virial) <= ftol * max(1.0, abs(mo.Virial)), (mg.Virial, mo.Virial)
    if mo.Options.Compute:
        for f in ("Potential", "Dispersion", "Coulomb"):
            a, b = getattr(mg.Energy, f), getattr(mo.Energy, f)
            assert abs(a - b) <= ftol * max(1.0, abs(b)), (f, a, b)


# ---------------------------------------------------------------- pair sets

def test_pairs_nist_lj(gpu, oracle):
    import hosts
    cfg = hosts.read_configuration(hosts.GOLDEN + "/NIST_lj_sample.mol")
    R, L = cfg["R"], cfg["L"]
    mg = pu.build(gpu, R, L, 3.0, 0.3, [lj()])
    mo = pu.build(oracle, R, L, 3.0, 0.3, [lj()])
    sg, so = pu.gpu_pairs(gpu, mg), pu.oracle_pairs(oracle, mo)
    assert sg == so
    assert sg == pu.brute_pairs(R, L, 3.3)


@pytest.mark.parametrize("case", ["fcc4000", "small_box_M5", "random_outside_box"])
def test_pairs_sets(gpu, oracle, case):
    if case == "fcc4000":
        R, L = pu.fcc(4000, 0.8442, jitter=0.05, seed=1)
        rc, skin = 2.5, 0.3
    elif case == "small_box_M5":      # 2L/(rc+skin) < 5: the grid is forced to 5 cells
        R, L = pu.fcc(256, 0.8442, jitter=0.1, seed=2)
        rc, skin = 2.5, 0.3
    else:                              # coordinates far outside [0, L)
        L = 9.0
        R = pu.random_box(300, L, seed=3)
        rc, skin = 2.0, 0.4
    mg = pu.build(gpu, R, L, rc, skin, [lj()])
    mo = pu.build(oracle, R, L, rc, skin, [lj()])
    sg, so = pu.gpu_pairs(gpu, mg), pu.oracle_pairs(oracle, mo)
    assert sg == so
    assert sg == pu.brute_pairs(R, L, rc + skin)


def test_pairs_exclusions_types(gpu, oracle):
    """ignore_pair, a non-interacting type pair (pair_none, no charges)."""
    R, L = pu.fcc(500, 0.8, jitter=0.05, seed=4)
    N = len(R)
    types = (np.arange(N) % 3 + 1).astype(np.int32)
    excl = [(1, 2), (5, 9), (10, 11), (100, 101), (3, 400)]
    models = [lj(), lj(1.2, 0.9), lambda lib: lib.EmDee_pair_none()]
    mg = pu.build(gpu, R, L, 2.5, 0.3, models, types=types, masses=[1.0, 2.0, 3.0], exclusions=excl)
    mo = pu.build(oracle, R, L, 2.5, 0.3, models, types=types, masses=[1.0, 2.0, 3.0], exclusions=excl)
    sg, so = pu.gpu_pairs(gpu, mg), pu.oracle_pairs(oracle, mo)
    assert sg == so
    # type 3 interacts with nobody (pair_none mixes to pair_none, no Coulomb)
    inter = [[1, 1, 0], [1, 1, 0], [0, 0, 0]]
    ex = set(excl) | {(j, i) for i, j in excl}
    assert sg == pu.brute_pairs(R, L, 2.8, exclude=ex, types=types, interact=inter)
    compare_state(gpu, oracle, mg, mo, N)


# ------------------------------------------------------------------ forces

MODS = {
    "plain": lambda m: lambda lib: m(lib),
    "shifted": lambda m: lambda lib: lib.EmDee_shifted(m(lib)),
    "shifted_force": lambda m: lambda lib: lib.EmDee_shifted_force(m(lib)),
    "smoothed": lambda m: lambda lib: lib.EmDee_smoothed(m(lib), 0.6),
    "shifted_smoothed": lambda m: lambda lib: lib.EmDee_shifted_smoothed(m(lib), 0.6),
    "square_smoothed": lambda m: lambda lib: lib.EmDee_square_smoothed(m(lib), 0.6),
    "shifted_square_smoothed": lambda m: lambda lib: lib.EmDee_shifted_square_smoothed(m(lib), 0.6),
}


@pytest.mark.parametrize("compute", [True, False])
@pytest.mark.parametrize("mod", list(MODS))
@pytest.mark.parametrize("model", ["lj", "softcore"])
def test_forces_models(gpu, oracle, model, mod, compute):
    R, L = pu.fcc(2048, 0.8442, jitter=0.08, seed=5)
    base = lj() if model == "lj" else (lambda lib: lib.EmDee_pair_softcore_cut(1.0, 1.0, 0.7))
    m = MODS[mod](base)
    mg = pu.build(gpu, R, L, 2.5, 0.3, [m], compute=compute)
    mo = pu.build(oracle, R, L, 2.5, 0.3, [m], compute=compute)
    compare_state(gpu, oracle, mg, mo, len(R))


def test_forces_mixing(gpu, oracle):
    """Lorentz-Berthelot mixing of lj/softcore (pair_softcore_cut.f90:139-163)
    plus an explicit cross pair that is not overridden by later mixing."""
    R, L = pu.fcc(1500, 0.8, jitter=0.05, seed=6)
    N = len(R)
    types = (np.arange(N) % 3 + 1).astype(np.int32)
    models = [lj(1.0, 1.0), lambda lib: lib.EmDee_pair_softcore_cut(0.8, 1.1, 0.5), lj(1.3, 0.95)]
    cross = {(1, 3): lambda lib: lib.EmDee_shifted_force(lib.EmDee_pair_lj_cut(0.5, 1.05))}
    mg = pu.build(gpu, R, L, 2.6, 0.3, models, types=types, masses=[1.0, 1.5, 2.0], cross=cross)
    mo = pu.build(oracle, R, L, 2.6, 0.3, models, types=types, masses=[1.0, 1.5, 2.0], cross=cross)
    compare_state(gpu, oracle, mg, mo, N)


COULS = {
    "cut": lambda lib: lib.EmDee_coul_cut(),
    "sf": lambda lib: lib.EmDee_coul_sf(),
    "damped": lambda lib: lib.EmDee_coul_damped(0.35),
    "damped_smoothed": lambda lib: lib.EmDee_coul_damped_smoothed(0.35, 0.5),
    "damped_square_smoothed": lambda lib: lib.EmDee_coul_damped_square_smoothed(0.35, 0.5),
    "square_smoothed": lambda lib: lib.EmDee_coul_square_smoothed(0.5),
    "shifted_square_smoothed": lambda lib: lib.EmDee_coul_shifted_square_smoothed(0.5),
    "shifted(damped)": lambda lib: lib.EmDee_shifted(lib.EmDee_coul_damped(0.35)),
    "none": lambda lib: lib.EmDee_coul_none(),
}


@pytest.mark.parametrize("compute", [True, False])
@pytest.mark.parametrize("coul", list(COULS))
def test_forces_coulomb(gpu, oracle, coul, compute):
    R, L = pu.fcc(1372, 0.8, jitter=0.05, seed=7)
    N = len(R)
    rng = np.random.default_rng(8)
    q = rng.choice([-0.5, 0.5], N)
    q[::7] = 0.0
    mg = pu.build(gpu, R, L, 3.0, 0.3, [lj()], charges=q, coul=COULS[coul], kcoul=1.3, compute=compute)
    mo = pu.build(oracle, R, L, 3.0, 0.3, [lj()], charges=q, coul=COULS[coul], kcoul=1.3, compute=compute)
    compare_state(gpu, oracle, mg, mo, N)


def test_dynamics_and_trigger(gpu, oracle):
    """50 MD steps: same rebuild instants (Builds) and matching energies."""
    R, L = pu.fcc(4000, 0.8442)
    outs = []
    for lib in (gpu, oracle):
        md = pu.build(lib, R, L, 2.5, 0.3, [lj()])
        lib.random_momenta(md, 1.5, True, 86245)
        for _ in range(50):
            lib.boost(md, 1.0, 0.0, 0.0025)
            lib.displace(md, 1.0, 0.0, 0.005)
            lib.boost(md, 1.0, 0.0, 0.0025)
        outs.append((md.Builds, md.Energy.Potential, md.Virial, md.Kinetic.Total,
                     lib.download(md, "coordinates", (len(R), 3))))
    (bg, ug, wg, kg, Rg), (bo, uo, wo, ko, Ro) = outs
    assert bg == bo
    assert abs(ug - uo) <= 1e-9 * abs(uo) and abs(wg - wo) <= 1e-8 * abs(wo) and abs(kg - ko) <= 1e-9 * abs(ko)
    assert np.max(np.abs(Rg - Ro)) < 1e-9


def test_trigger_quirk(gpu, oracle):
    """maximum_approach_sq (neighbor_lists.f90:41-59) uses the running maximum
    before the first atom holding the maximum, not the true second largest.
    Displacements are crafted so that the reference rule does NOT rebuild while
    the classic two-largest rule would; both libraries must agree."""
    R, L = pu.fcc(864, 0.8442, jitter=0.02, seed=9)
    N = len(R)
    skin = 0.3
    builds = []
    for lib in (gpu, oracle):
        md = pu.build(lib, R, L, 2.5, skin, [lj()])
        b0 = md.Builds
        D = np.zeros_like(R)
        D[0, 0] = 0.05          # atom 1: small
        D[10, 1] = 0.20         # first atom holding the maximum
        D[500, 2] = 0.12        # true second largest, after it
        lib.upload(md, "coordinates", R + D)
        lib.compute_forces(md)
        b1 = md.Builds
        D[10, 1] = 0.26         # now 0.26 + 0.05 > 0.3 fires
        lib.upload(md, "coordinates", R + D)
        lib.compute_forces(md)
        builds.append((b1 - b0, md.Builds - b0))
    # reference rule: (0.2+0.05)^2 < 0.09 -> no build; (0.26+0.05)^2 > 0.09 -> build
    assert builds[0] == builds[1] == (0, 1)


def test_mixed_precision_error(gpu, oracle):
    R, L = pu.fcc(8788, 0.8442, jitter=0.05, seed=10)
    mg = pu.build(gpu, R, L, 2.5, 0.3, [lj()])
    gpu.EmDeeGPU_set_precision(mg, 1)
    mo = pu.build(oracle, R, L, 2.5, 0.3, [lj()])
    Fg = gpu.download(mg, "forces", (len(R), 3))
    Fo = oracle.download(mo, "forces", (len(R), 3))
    err = np.linalg.norm(Fg - Fo) / np.linalg.norm(Fo)
    print("mixed-form force RMS error", err)
    assert err < 1e-5
    assert abs(mg.Energy.Potential - mo.Energy.Potential) < 1e-6 * abs(mo.Energy.Potential)


@pytest.mark.parametrize("natoms", [1, 2, 7])
@pytest.mark.parametrize("mixed", [0, 1])
def test_tiny_and_dilute_systems(gpu, oracle, natoms, mixed):
    """Edge cases of the list and pair kernels: one atom (an empty list),
    two atoms, and a dilute box where most atoms have no neighbour; then a
    few MD steps with identical rebuild instants."""
    L = 12.0
    R = pu.random_box(natoms, L, seed=20 + natoms, min_dist=1.0)
    mg = pu.build(gpu, R, L, 2.5, 0.3, [lj()])
    mo = pu.build(oracle, R, L, 2.5, 0.3, [lj()])
    gpu.EmDeeGPU_set_precision(mg, mixed)
    assert pu.gpu_pairs(gpu, mg) == pu.oracle_pairs(oracle, mo) == pu.brute_pairs(R, L, 2.8)
    compare_state(gpu, oracle, mg, mo, natoms, ftol=1e-6 if mixed else FTOL)
    for lib, md in ((gpu, mg), (oracle, mo)):
        lib.random_momenta(md, 1.0, False, 5)    # no momentum adjustment: one atom has no free dof
        bench_like_steps(lib, md, 20, 0.005)
    assert mg.Builds == mo.Builds
    tol = 1e-6 if mixed else 1e-10
    assert np.max(np.abs(gpu.download(mg, "coordinates", (natoms, 3)) -
                         oracle.download(mo, "coordinates", (natoms, 3)))) < tol


def bench_like_steps(lib, md, n, dt):
    for _ in range(n):
        md.Options.Compute = True
        lib.boost(md, 1.0, 0.0, 0.5 * dt)
        lib.displace(md, 1.0, 0.0, dt)
        lib.boost(md, 1.0, 0.0, 0.5 * dt)
