"""Rigid 3-site water set-ups used by the rigid-body and Coulomb tests.

`spce_fixture_system` follows test/test_rigid_body_exact.f90:36-50 on the
NIST SPC/E fixture (the reference's spce_sample.inp names a missing file;
NIST_spce_sample.mol is the fixture that exists): bodies = molecule ids,
shifted-force LJ for O, pair_none for H, Coulomb with kCoul of
spce_sample.inp, Da-A-fs units.
"""
import ctypes

import numpy as np

import hosts

KB = 8.3144621455e-7           # Da*A^2/(fs^2*K), spce_sample.inp
KCOUL = 0.13893545755135628    # Da*A^3/(fs*e)^2
MVV2E = 2390.057364            # (kcal/mol)/(Da*A^2/fs^2)


def spce_fixture_system(lib, rc=10.0, skin=2.0, coul="sf", mode=0, threads=1, kspace=None):
    cfg = hosts.read_configuration(hosts.GOLDEN + "/NIST_spce_sample.mol")
    eps = cfg["epsilon"] / MVV2E
    md = lib.system(threads, 1, rc, skin, cfg["N"], cfg["types"], cfg["mass"], cfg["molecule"])
    for i in range(cfg["ntypes"]):
        if eps[i] == 0.0:
            pair = lib.EmDee_pair_none()
        else:
            pair = lib.EmDee_shifted_force(lib.EmDee_pair_lj_cut(eps[i], cfg["sigma"][i]))
        lib.EmDee_set_pair_model(md, i + 1, i + 1, pair, KCOUL)
    if coul == "sf":
        lib.EmDee_set_coul_model(md, lib.EmDee_coul_sf())
    elif coul == "damped_smoothed":
        lib.EmDee_set_coul_model(md, lib.EmDee_coul_damped_smoothed(0.2, 1.0))
    elif coul == "long":
        lib.EmDee_set_coul_model(md, lib.EmDee_coul_long())
        lib.EmDee_set_kspace_model(md, lib.EmDee_kspace_ewald(kspace))
    md.Options.RotationMode = mode
    lib.upload(md, "charges", cfg["Q"])
    lib.upload(md, "coordinates", cfg["R"])
    lib.upload(md, "box", np.array([cfg["L"]]))
    return md, cfg


def run(lib, md, nsteps, dt):
    for _ in range(nsteps):
        lib.boost(md, 1.0, 0.0, 0.5 * dt)
        lib.displace(md, 1.0, 0.0, dt)
        lib.boost(md, 1.0, 0.0, 0.5 * dt)


def two_layer_fixture(lib, inner):
    """The NIST SPC/E fixture with two layers: layer 1 LJ + coul_long +
    kspace_ewald(1e-5); layer 2 shifted-force LJ + coul_damped(0.25); with
    `inner`, layer 2 uses an inner cutoff of 6 A (layer_based_parameters)."""
    cfg = hosts.read_configuration(hosts.GOLDEN + "/NIST_spce_sample.mol")
    md = lib.system(1, 2, 9.0, 1.0, cfg["N"], cfg["types"], cfg["mass"], cfg["molecule"])
    if inner:
        apply = (ctypes.c_int * 2)(0, 1)
        bonded = (ctypes.c_int * 2)(1, 1)
        lib.EmDee_layer_based_parameters(md, 6.0, apply, bonded)
    eps = cfg["epsilon"] / MVV2E
    kc = (ctypes.c_double * 2)(KCOUL, KCOUL)
    for i in range(cfg["ntypes"]):
        if eps[i] == 0.0:
            models = (ctypes.c_void_p * 2)(lib.EmDee_pair_none(), lib.EmDee_pair_none())
        else:
            models = (ctypes.c_void_p * 2)(lib.EmDee_pair_lj_cut(eps[i], cfg["sigma"][i]),
                                           lib.EmDee_shifted_force(lib.EmDee_pair_lj_cut(eps[i], cfg["sigma"][i])))
        lib.EmDee_set_pair_multimodel(md, i + 1, i + 1, models, kc)
    lib.EmDee_set_coul_multimodel(md, (ctypes.c_void_p * 2)(lib.EmDee_coul_long(), lib.EmDee_coul_damped(0.25)))
    lib.EmDee_set_kspace_model(md, lib.EmDee_kspace_ewald(1e-5))
    lib.upload(md, "charges", cfg["Q"])
    lib.upload(md, "coordinates", cfg["R"])
    lib.upload(md, "box", np.array([cfg["L"]]))
    return md, cfg
