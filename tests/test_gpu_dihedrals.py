"""Harmonic dihedrals on the GPU (bonded.hip k_dihedrals) against the oracle:
a fluid of 4-site chains with LJ between chains, harmonic bonds, angles and
dihedrals; forces, energy terms and virial on identical coordinates, then
after 50 velocity-Verlet steps (same rebuilds)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def chains(lib, nchain=125, seed=5):
    rng = np.random.default_rng(seed)
    m = int(round(nchain ** (1 / 3)))
    L = m * 3.2
    R = []
    for c in range(nchain):
        o = (np.array([c % m, (c // m) % m, c // (m * m)]) + 0.2) * (L / m)
        base = np.array([[0, 0, 0], [0.8, 0.55, 0], [1.6, 0.0, 0.2], [2.4, 0.5, 0.7]])
        R.append(o + base + 0.05 * rng.standard_normal((4, 3)))
    R = np.concatenate(R)
    N = len(R)
    md = lib.system(1, 1, 2.5, 0.3, N, None, None, None)
    lib.EmDee_set_pair_model(md, 1, 1, lib.EmDee_pair_lj_cut(0.2, 0.8), 0.0)
    for c in range(nchain):
        a = 4 * c + 1
        for i in range(3):
            lib.EmDee_add_bond(md, a + i, a + i + 1, lib.EmDee_bond_harmonic(100.0, 1.0))
        for i in range(2):
            lib.EmDee_add_angle(md, a + i, a + i + 1, a + i + 2, lib.EmDee_angle_harmonic(20.0, 1.9))
        lib.EmDee_add_dihedral(md, a, a + 1, a + 2, a + 3, lib.EmDee_dihedral_harmonic(3.0, 0.4))
    lib.upload(md, "box", np.array([L]))
    lib.upload(md, "coordinates", R)
    lib.random_momenta(md, 0.8, True, 7)
    return md, N


def test_dihedral_chains_match_oracle(gpu, oracle):
    mg, N = chains(gpu)
    mo, _ = chains(oracle)
    for stage in range(2):
        if stage:
            for lib, md in ((gpu, mg), (oracle, mo)):
                for _ in range(50):
                    lib.boost(md, 1.0, 0.0, 0.001)
                    lib.displace(md, 1.0, 0.0, 0.002)
                    lib.boost(md, 1.0, 0.0, 0.001)
        gpu.compute_forces(mg)
        oracle.compute_forces(mo)
        Fg, Fo = gpu.download(mg, "forces", (N, 3)), oracle.download(mo, "forces", (N, 3))
        print(f"stage {stage}: dihedral E {mg.Energy.Dihedral:.12f} / {mo.Energy.Dihedral:.12f}, "
              f"max force diff {np.max(np.abs(Fg - Fo)):.2e}, builds {mg.Builds}/{mo.Builds}", flush=True)
        assert mo.Energy.Dihedral > 0.0
        assert np.max(np.abs(Fg - Fo)) <= 1e-11 * np.max(np.abs(Fo))
        for f in ("Potential", "Dispersion", "Bond", "Angle", "Dihedral"):
            assert getattr(mg.Energy, f) == pytest.approx(getattr(mo.Energy, f), rel=1e-11, abs=1e-12)
        assert mg.Virial == pytest.approx(mo.Virial, rel=1e-10, abs=1e-10)
        assert mg.Builds == mo.Builds
