"""Mixed form (FP32 pair arithmetic, FP64 accumulation; EmDeeGPU_set_precision)
against the FP64 oracle on every model family: force RMS error relative to
the force norm, energy and virial errors, per configuration.  The mixed form
keeps the FP64 form's pair set (cutoff-band pairs are decided in FP64), so
the neighbor-list builds match the oracle's too."""
import numpy as np
import pytest

import hosts
import ionic
import water

pytestmark = pytest.mark.gpu

FORCE_RMS = 1e-5      # relative RMS force error bound of the mixed form


def errors(gpu, oracle, mg, mo, N):
    Fg, Fo = gpu.download(mg, "forces", (N, 3)), oracle.download(mo, "forces", (N, 3))
    ferr = np.linalg.norm(Fg - Fo) / np.linalg.norm(Fo)
    eerr = abs(mg.Energy.Potential - mo.Energy.Potential) / max(abs(mo.Energy.Potential), 1e-300)
    werr = abs(mg.Virial - mo.Virial) / max(abs(mo.Virial), 1e-300)
    return ferr, eerr, werr


@pytest.mark.parametrize("name", ["lj_cut", "lj_sf", "lj_smoothed", "lj_shifted_smoothed"])
def test_lj_models(gpu, oracle, name):
    mg, cfg = hosts.lj_sample_system(gpu, name)
    gpu.EmDeeGPU_set_precision(mg, 1)
    gpu.compute_forces(mg)
    mo, _ = hosts.lj_sample_system(oracle, name)
    ferr, eerr, werr = errors(gpu, oracle, mg, mo, cfg["N"])
    dW = abs(mg.Virial - mo.Virial) / cfg["N"]
    print(f"mixed {name}: force rms {ferr:.2e} energy {eerr:.2e} virial {werr:.2e} (per atom {dW:.1e} eps)")
    # the LJ virial is a sum of large terms of both signs, so its relative
    # error is amplified by cancellation; the per-atom error is what enters
    # the pressure beside the kinetic term (order kT ~ eps per atom)
    assert ferr < FORCE_RMS and eerr < 1e-6 and dW < 2e-5


@pytest.mark.parametrize("coul", ["sf", "damped_smoothed", "long"])
def test_rigid_water(gpu, oracle, coul):
    kw = dict(coul=coul, kspace=1e-5 if coul == "long" else None)
    mg, cfg = water.spce_fixture_system(gpu, **kw)
    gpu.EmDeeGPU_set_precision(mg, 1)
    gpu.compute_forces(mg)
    mo, _ = water.spce_fixture_system(oracle, **kw)
    ferr, eerr, werr = errors(gpu, oracle, mg, mo, cfg["N"])
    print(f"mixed water/{coul}: force rms {ferr:.2e} energy {eerr:.2e} virial {werr:.2e}")
    assert ferr < FORCE_RMS and eerr < 1e-6 and werr < 1e-5
    assert mg.BodyVirial == pytest.approx(mo.BodyVirial, rel=1e-4)


def test_rock_salt(gpu, oracle):
    kw = dict(ncell=6, rc=4.0, skin=0.2, accuracy=1e-6, jitter=0.2)
    mg, cfg = ionic.nacl_system(gpu, **kw)
    gpu.EmDeeGPU_set_precision(mg, 1)
    gpu.compute_forces(mg)
    mo, _ = ionic.nacl_system(oracle, **kw)
    ferr, eerr, werr = errors(gpu, oracle, mg, mo, cfg["N"])
    print(f"mixed rock salt: force rms {ferr:.2e} energy {eerr:.2e} virial {werr:.2e}")
    assert ferr < FORCE_RMS and eerr < 1e-6


def test_softcore(gpu, oracle):
    cfg = hosts.read_configuration(hosts.GOLDEN + "/NIST_lj_sample.mol")

    def make(lib, mixed):
        md = lib.system(1, 1, 3.0, 0.3, cfg["N"], cfg["types"], cfg["mass"], None)
        lib.EmDee_set_pair_model(md, 1, 1, lib.EmDee_pair_softcore_cut(1.0, 1.0, 0.6), 0.0)
        if mixed:
            lib.EmDeeGPU_set_precision(md, 1)
        lib.upload(md, "box", np.array([cfg["L"]]))
        lib.upload(md, "coordinates", cfg["R"])
        return md
    mo, mg = make(oracle, False), make(gpu, True)
    ferr, eerr, werr = errors(gpu, oracle, mg, mo, cfg["N"])
    print(f"mixed softcore: force rms {ferr:.2e} energy {eerr:.2e} virial {werr:.2e}")
    assert ferr < FORCE_RMS and eerr < 1e-6


def test_same_pair_set_as_fp64(gpu):
    """Dynamics in the mixed form rebuild on the same schedule and keep the
    interacting pair set of the FP64 form on identical coordinates."""
    mg, cfg = hosts.lj_sample_system(gpu, "lj_cut")
    gpu.EmDeeGPU_set_precision(mg, 1)
    for _ in range(40):
        gpu.verlet_step(mg, 0.005)
    R = gpu.download(mg, "coordinates", (cfg["N"], 3))
    E_mixed = mg.Energy.Potential
    gpu.EmDeeGPU_set_precision(mg, 0)
    gpu.compute_forces(mg)
    assert mg.Energy.Potential == pytest.approx(E_mixed, rel=1e-6)
    assert np.array_equal(R, gpu.download(mg, "coordinates", (cfg["N"], 3)))


def fcc_box(nc, density, jitter, seed):
    """4 nc^3 atoms on an fcc lattice that spans a cubic box at the given
    density, each displaced by a seeded Gaussian of deviation `jitter`."""
    N = 4 * nc ** 3
    L = (N / density) ** (1.0 / 3.0)
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(*(np.arange(nc),) * 3, indexing="ij"), -1).reshape(-1, 1, 3)
    R = ((g + base) * (L / nc)).reshape(-1, 3)
    R += np.random.default_rng(seed).normal(0.0, jitter, R.shape)
    return R, L


@pytest.mark.parametrize("mixed,shifted", [(True, False), (True, True), (False, False)])
def test_large_lj_fast_path(gpu, oracle, mixed, shifted):
    """131,072 atoms of one type, lj_cut (plain and shifted): the size from
    which both forms run their dedicated LJ kernels (in the mixed form,
    blocks of twelve entries with cutoff-band entries decided in FP64) and
    pair evaluations are queued ahead of the rebuild trigger's answer.
    Forces, energy, virial, kinetic energy and rebuild count against the
    oracle, before and after twenty steps with list rebuilds."""
    R, L = fcc_box(32, 0.8442, 0.08, seed=17)
    N = len(R)
    tol = FORCE_RMS if mixed else 1e-9

    def make(lib, gpu_mixed):
        md = lib.system(1, 1, 2.5, 0.3, N, None, None, None)
        model = lib.EmDee_pair_lj_cut(1.0, 1.0)
        if shifted:
            model = lib.EmDee_shifted(model)
        lib.EmDee_set_pair_model(md, 1, 1, model, 0.0)
        if gpu_mixed:
            lib.EmDeeGPU_set_precision(md, 1)
        lib.upload(md, "box", np.array([L]))
        lib.upload(md, "coordinates", R)
        lib.random_momenta(md, 1.0, True, 5)
        return md
    mg, mo = make(gpu, mixed), make(oracle, False)
    for stage in range(2):
        if stage:
            for _ in range(20):
                for lib, md in ((gpu, mg), (oracle, mo)):
                    lib.boost(md, 1.0, 0.0, 0.0025)
                    lib.displace(md, 1.0, 0.0, 0.005)
                    lib.boost(md, 1.0, 0.0, 0.0025)
        gpu.compute_forces(mg)
        oracle.compute_forces(mo)
        ferr, eerr, werr = errors(gpu, oracle, mg, mo, N)
        dW = abs(mg.Virial - mo.Virial) / N
        kerr = abs(mg.Kinetic.Total - mo.Kinetic.Total) / mo.Kinetic.Total
        print(f"LJ N={N} mixed={mixed} shifted={shifted} stage {stage}: force rms {ferr:.2e} energy {eerr:.2e} "
              f"virial per atom {dW:.1e} kinetic {kerr:.1e} builds {mg.Builds}/{mo.Builds}")
        assert ferr < tol and eerr < 1e-6 and dW < 2e-5 and kerr < 1e-6
        assert mg.Builds == mo.Builds
