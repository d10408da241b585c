"""Ionic set-ups for the Ewald tests: a rock-salt crystal (NaCl structure)
and flexible 3-site water (the NIST SPC/E fixture held together by harmonic
bonds and angles instead of rigid bodies)."""
import numpy as np

import hosts
import water

MADELUNG_NACL = 1.747564594633182   # rock salt, per ion pair, nearest-neighbour distance 1


def rock_salt(ncell, r0=1.0):
    """Positions and charges of ncell^3 conventional cells (8 ions each) of
    the rock-salt lattice with nearest-neighbour distance r0."""
    g = np.arange(2 * ncell)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    R = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).astype(float) * r0
    Q = np.where((X + Y + Z).ravel() % 2 == 0, 1.0, -1.0)
    return R, Q, 2 * ncell * r0


def nacl_system(lib, ncell=4, rc=3.5, skin=0.3, accuracy=1e-6, jitter=0.0, seed=7, kcoul=1.0, layers=1):
    R, Q, L = rock_salt(ncell)
    if jitter:
        R = R + np.random.default_rng(seed).uniform(-jitter, jitter, R.shape)
    N = len(Q)
    types = np.where(Q > 0, 1, 2).astype(np.int32)
    md = lib.system(1, layers, rc, skin, N, types, np.array([22.99, 35.45]), None)
    for l in range(1, layers + 1):
        for i in (1, 2):
            lib.EmDee_set_pair_model(md, i, i, lib.EmDee_pair_none(), kcoul)
        lib.EmDee_set_coul_model(md, lib.EmDee_coul_long())
    lib.EmDee_set_kspace_model(md, lib.EmDee_kspace_ewald(accuracy))
    lib.upload(md, "charges", Q)
    lib.upload(md, "coordinates", R)
    lib.upload(md, "box", np.array([L]))
    return md, dict(R=R, Q=Q, L=L, N=N, types=types)


def flexible_water(lib, coul="sf", rc=9.0, skin=1.0, accuracy=1e-5, kb=1.0, ka=0.1, r0=0.97, theta0=104.5):
    """NIST SPC/E configuration with O-H bonds and the H-O-H angle as
    harmonic terms (no rigid bodies); kb, ka in Da A^2/fs^2 per A^2 / rad^2.
    The fixture holds SPC/E geometry (1 A, 109.47 deg); r0 and theta0 differ
    from it so that the bonded energies and forces start away from zero."""
    cfg = hosts.read_configuration(hosts.GOLDEN + "/NIST_spce_sample.mol")
    eps = cfg["epsilon"] / water.MVV2E
    md = lib.system(1, 1, rc, skin, cfg["N"], cfg["types"], cfg["mass"], None)
    for i in range(cfg["ntypes"]):
        if eps[i] == 0.0:
            pair = lib.EmDee_pair_none()
        else:
            pair = lib.EmDee_shifted_force(lib.EmDee_pair_lj_cut(eps[i], cfg["sigma"][i]))
        lib.EmDee_set_pair_model(md, i + 1, i + 1, pair, water.KCOUL)
    if coul == "sf":
        lib.EmDee_set_coul_model(md, lib.EmDee_coul_sf())
    else:
        lib.EmDee_set_coul_model(md, lib.EmDee_coul_long())
        lib.EmDee_set_kspace_model(md, lib.EmDee_kspace_ewald(accuracy))
    mol = cfg["molecule"]
    types = cfg["types"]
    for m in np.unique(mol):
        idx = np.where(mol == m)[0] + 1
        o = [i for i in idx if types[i - 1] == 1][0]
        hs = [i for i in idx if types[i - 1] != 1]
        for h in hs:
            lib.EmDee_add_bond(md, int(o), int(h), lib.EmDee_bond_harmonic(kb, r0))
        lib.EmDee_add_angle(md, int(hs[0]), int(o), int(hs[1]), lib.EmDee_angle_harmonic(ka, np.deg2rad(theta0)))
    lib.upload(md, "charges", cfg["Q"])
    lib.upload(md, "coordinates", cfg["R"])
    lib.upload(md, "box", np.array([cfg["L"]]))
    return md, cfg
