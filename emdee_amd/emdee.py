"""ctypes binding of the EmDee C ABI declared in include/emdee.h.

The binding mirrors the reference's Fortran interface module
(src/emdee_header.f03:20-323): the same entry points, argument meaning and
1-based indices, so host code written in Python reads like the reference's
test programs (test/test_pair_lj_cut.f90).  `EmDeeLibrary(path)` wraps any
shared library exporting that ABI.
"""
import ctypes
from ctypes import (POINTER, Structure, c_bool, c_char_p, c_double, c_int,
                    c_void_p)

import numpy as np


class tTime(Structure):
    _fields_ = [("Pair", c_double), ("Motion", c_double),
                ("Neighbor", c_double), ("Total", c_double)]


class tEnergy(Structure):
    _fields_ = [("Potential", c_double), ("Dispersion", c_double),
                ("Coulomb", c_double), ("Bond", c_double), ("Angle", c_double),
                ("Dihedral", c_double), ("ShadowPotential", c_double),
                ("UpToDate", c_bool)]


class tKinetic(Structure):
    _fields_ = [("Total", c_double), ("TransPart", c_double * 3),
                ("Rotational", c_double), ("RotPart", c_double * 3),
                ("ShadowKinetic", c_double), ("ShadowRotational", c_double),
                ("UpToDate", c_bool)]


class tOpts(Structure):
    _fields_ = [("Translate", c_bool), ("Rotate", c_bool),
                ("RotationMode", c_int), ("AutoForceCompute", c_bool),
                ("AutoBodyUpdate", c_bool), ("Compute", c_bool)]


class tEmDee(Structure):
    """Layout of src/emdee_header.f03:70-80 (240 bytes)."""
    _fields_ = [("Builds", c_int), ("Time", tTime), ("Energy", tEnergy),
                ("Kinetic", tKinetic), ("Virial", c_double),
                ("BodyVirial", c_double), ("DoF", c_int), ("RotDoF", c_int),
                ("Data", c_void_p), ("Options", tOpts)]


assert ctypes.sizeof(tEmDee) == 240

_dp = POINTER(c_double)
_ip = POINTER(c_int)

# name -> (restype, argtypes)
SIGNATURES = {
    "EmDee_system": (tEmDee, [c_int, c_int, c_double, c_double, c_int, _ip, _dp, _ip]),
    "EmDee_memory_address": (c_void_p, [tEmDee, c_char_p]),
    "EmDee_share_phase_space": (None, [tEmDee, POINTER(tEmDee)]),
    "EmDee_layer_based_parameters": (None, [tEmDee, c_double, _ip, _ip]),
    "EmDee_set_pair_model": (None, [tEmDee, c_int, c_int, c_void_p, c_double]),
    "EmDee_set_pair_multimodel": (None, [tEmDee, c_int, c_int, POINTER(c_void_p), _dp]),
    "EmDee_set_kspace_model": (None, [tEmDee, c_void_p]),
    "EmDee_set_coul_model": (None, [tEmDee, c_void_p]),
    "EmDee_set_coul_multimodel": (None, [tEmDee, POINTER(c_void_p)]),
    "EmDee_ignore_pair": (None, [tEmDee, c_int, c_int]),
    "EmDee_add_bond": (None, [tEmDee, c_int, c_int, c_void_p]),
    "EmDee_add_angle": (None, [tEmDee, c_int, c_int, c_int, c_void_p]),
    "EmDee_add_dihedral": (None, [tEmDee, c_int, c_int, c_int, c_int, c_void_p]),
    "EmDee_download": (None, [tEmDee, c_char_p, _dp]),
    "EmDee_upload": (None, [POINTER(tEmDee), c_char_p, _dp]),
    "EmDee_switch_model_layer": (None, [POINTER(tEmDee), c_int]),
    "EmDee_random_momenta": (None, [POINTER(tEmDee), c_double, c_bool, c_int]),
    "EmDee_boost": (None, [POINTER(tEmDee), c_double, c_double, c_double]),
    "EmDee_displace": (None, [POINTER(tEmDee), c_double, c_double, c_double]),
    "EmDee_move": (None, [POINTER(tEmDee), c_double, c_double, c_double]),
    "EmDee_verlet_step": (None, [POINTER(tEmDee), c_double]),
    "EmDee_compute_forces": (None, [POINTER(tEmDee)]),
    "EmDee_rdf": (None, [tEmDee, c_int, c_double, c_int, _ip, _ip, _dp]),
    "EmDee_shifted": (c_void_p, [c_void_p]),
    "EmDee_shifted_force": (c_void_p, [c_void_p]),
    "EmDee_smoothed": (c_void_p, [c_void_p, c_double]),
    "EmDee_shifted_smoothed": (c_void_p, [c_void_p, c_double]),
    "EmDee_square_smoothed": (c_void_p, [c_void_p, c_double]),
    "EmDee_shifted_square_smoothed": (c_void_p, [c_void_p, c_double]),
    "EmDee_pair_none": (c_void_p, []),
    "EmDee_coul_none": (c_void_p, []),
    "EmDee_bond_none": (c_void_p, []),
    "EmDee_angle_none": (c_void_p, []),
    "EmDee_dihedral_none": (c_void_p, []),
    "EmDee_pair_lj_cut": (c_void_p, [c_double, c_double]),
    "EmDee_pair_softcore_cut": (c_void_p, [c_double, c_double, c_double]),
    "EmDee_coul_cut": (c_void_p, []),
    "EmDee_coul_sf": (c_void_p, []),
    "EmDee_coul_damped": (c_void_p, [c_double]),
    "EmDee_coul_long": (c_void_p, []),
    "EmDee_coul_damped_smoothed": (c_void_p, [c_double, c_double]),
    "EmDee_coul_damped_square_smoothed": (c_void_p, [c_double, c_double]),
    "EmDee_coul_square_smoothed": (c_void_p, [c_double]),
    "EmDee_coul_shifted_square_smoothed": (c_void_p, [c_double]),
    "EmDee_bond_harmonic": (c_void_p, [c_double, c_double]),
    "EmDee_angle_harmonic": (c_void_p, [c_double, c_double]),
    "EmDee_dihedral_harmonic": (c_void_p, [c_double, c_double]),
    "EmDee_kspace_ewald": (c_void_p, [c_double]),
}


# Extensions of the HIP build (include/emdee_gpu.h); bound when present.
GPU_SIGNATURES = {
    "EmDeeGPU_set_precision": (None, [tEmDee, c_int]),
    "EmDeeGPU_get_precision": (c_int, [tEmDee]),
    "EmDeeGPU_neighbor_pairs": (c_int, [tEmDee, c_int, _ip]),
    "EmDeeGPU_profile": (None, [tEmDee, c_int]),
    "EmDeeGPU_kernel_stats": (None, [tEmDee, _dp]),
    "EmDeeGPU_synchronize": (None, [tEmDee]),
    "EmDeeGPU_destroy": (None, [POINTER(tEmDee)]),
    "EmDeeHMC_create": (c_void_p, [c_int, _ip, _dp, _dp, _dp, c_double, c_double, c_int, ctypes.c_ulonglong, c_int]),
    "EmDeeHMC_tune": (c_int, [c_void_p]),
    "EmDeeHMC_variant": (None, [c_void_p, _ip, _ip, _dp]),
    "EmDeeHMC_run": (None, [c_void_p, c_int, c_int]),
    "EmDeeHMC_synchronize": (None, [c_void_p]),
    "EmDeeHMC_get": (None, [c_void_p, _dp, _dp, POINTER(ctypes.c_longlong), _dp]),
    "EmDeeHMC_history": (c_int, [c_void_p, c_int, _dp]),
    "EmDeeHMC_destroy": (None, [c_void_p]),
    "EmDeeHMC_momenta": (None, [ctypes.c_ulonglong, c_int, c_int, c_int, c_double, _dp]),
    "EmDeeHMC_uniform": (c_double, [ctypes.c_ulonglong, c_int, c_int]),
    "EmDeeGPU_atom_energies": (None, [POINTER(tEmDee), _dp]),
    "EmDeeSurrogate_create": (c_void_p, [c_int, c_double, ctypes.c_ulonglong]),
    "EmDeeGPU_surrogate_sample": (None, [tEmDee, c_void_p, _dp]),
    "EmDeeSurrogate_rows": (c_int, [c_void_p]),
    "EmDeeSurrogate_train": (c_double, [c_void_p, c_int, c_int, c_double]),
    "EmDeeSurrogate_param_count": (c_int, [c_void_p]),
    "EmDeeSurrogate_get_params": (None, [c_void_p, POINTER(ctypes.c_float), POINTER(ctypes.c_float)]),
    "EmDeeSurrogate_set_params": (None, [c_void_p, POINTER(ctypes.c_float), POINTER(ctypes.c_float)]),
    "EmDeeSurrogate_destroy": (None, [c_void_p]),
    "EmDeeGPU_attach_surrogate": (None, [tEmDee, c_void_p, c_int]),
    "EmDeeGPU_rdf_noalloc": (None, [tEmDee, c_int, c_double, c_int, _ip, _ip, _dp]),
    "EmDeeSurrogate_set_long_range": (None, [c_void_p, c_double]),
    "EmDeeSurrogate_long_range_vectors": (c_int, [c_void_p]),
}


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


class EmDeeLibrary:
    """Thin wrapper over one shared library that exports the EmDee ABI."""

    def __init__(self, path):
        self.path = str(path)
        self.lib = ctypes.CDLL(self.path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(self.lib, name)
            f.restype = res
            f.argtypes = args

        self.has_gpu_extensions = hasattr(self.lib, "EmDeeGPU_set_precision")
        if self.has_gpu_extensions:
            for name, (res, args) in GPU_SIGNATURES.items():
                f = getattr(self.lib, name)
                f.restype = res
                f.argtypes = args

    def __getattr__(self, name):
        if name.startswith("EmDee"):
            return getattr(self.lib, name)
        raise AttributeError(name)

    # -- convenience wrappers with numpy arrays ---------------------------
    def system(self, threads, layers, rc, skin, N, types=None, masses=None, bodies=None):
        keep = []
        def ip(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.int32)
            keep.append(a)
            return a.ctypes.data_as(_ip)
        def dp(a):
            if a is None:
                return None
            a = _f64(a)
            keep.append(a)
            return a.ctypes.data_as(_dp)
        return self.lib.EmDee_system(threads, layers, rc, skin, N, ip(types), dp(masses), ip(bodies))

    def upload(self, md, option, array):
        a = _f64(array)
        self.lib.EmDee_upload(ctypes.byref(md), option.encode(), a.ctypes.data_as(_dp))

    def download(self, md, option, shape):
        out = np.zeros(shape, dtype=np.float64)
        self.lib.EmDee_download(md, option.encode(), out.ctypes.data_as(_dp))
        return out

    def random_momenta(self, md, kT, adjust, seed):
        self.lib.EmDee_random_momenta(ctypes.byref(md), kT, adjust, seed)

    def boost(self, md, lam, alpha, dt):
        self.lib.EmDee_boost(ctypes.byref(md), lam, alpha, dt)

    def displace(self, md, lam, alpha, dt):
        self.lib.EmDee_displace(ctypes.byref(md), lam, alpha, dt)

    def compute_forces(self, md):
        self.lib.EmDee_compute_forces(ctypes.byref(md))

    def verlet_step(self, md, dt):
        self.lib.EmDee_verlet_step(ctypes.byref(md), dt)

    def switch_model_layer(self, md, layer):
        self.lib.EmDee_switch_model_layer(ctypes.byref(md), layer)

    def neighbor_pairs(self, md):
        """Unordered Verlet-list pair set as an (n, 2) array of 1-based i < j."""
        n = self.lib.EmDeeGPU_neighbor_pairs(md, 0, None)
        out = np.zeros((max(n, 1), 2), dtype=np.int32)
        self.lib.EmDeeGPU_neighbor_pairs(md, n, out.ctypes.data_as(_ip))
        return out[:n]

    def atom_energies(self, md, N):
        e = np.zeros(N)
        self.lib.EmDeeGPU_atom_energies(ctypes.byref(md), e.ctypes.data_as(_dp))
        return e

    def kernel_stats(self, md):
        out = np.zeros(8)
        self.lib.EmDeeGPU_kernel_stats(md, out.ctypes.data_as(_dp))
        return out
