"""ctypes binding of libvortexhip.so (include/vortexhip.h).  The library must be built (vphip.build); there is no
fallback: without it every entry point raises."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBPATH = os.path.join(PKG, "lib", "libvortexhip.so")

VP_OK, VP_E_ARGUMENT, VP_E_DIMENSION, VP_E_DEVICE, VP_E_SINGULAR, VP_E_CONVERGENCE = 0, -1, -2, -3, -4, -5


class AssemblyOptionsC(C.Structure):
    _fields_ = [("test", C.c_int), ("far_ratio", C.c_double), ("far_points", C.c_int), ("near_points", C.c_int),
                ("near_ratio", C.c_double), ("near_depth", C.c_int), ("touch_levels", C.c_int),
                ("touch_ratio", C.c_double), ("touch_order", C.c_int), ("corner_angle", C.c_double),
                ("contact_rel", C.c_double), ("kernel_far", C.c_double), ("precision", C.c_int),
                ("quad_touch_levels", C.c_int)]


class MixedMeshC(C.Structure):
    """vp_mixed_mesh"""
    _fields_ = [("nn", C.c_int), ("x", C.POINTER(C.c_double)), ("nquad", C.c_int), ("quad_nodes", C.POINTER(C.c_int)),
                ("quad_slots", C.POINTER(C.c_int)), ("ntri", C.c_int), ("tri_nodes", C.POINTER(C.c_int)),
                ("tri_slots", C.POINTER(C.c_int)), ("nslot", C.c_int), ("ntouch", C.c_int),
                ("touch", C.POINTER(C.c_int))]


class SurfaceC(C.Structure):
    _fields_ = [("nn", C.c_int), ("x", C.POINTER(C.c_double)), ("nt", C.c_int), ("tris", C.POINTER(C.c_int)),
                ("corner_slot", C.POINTER(C.c_int)), ("nslot", C.c_int), ("ntouch", C.c_int),
                ("touch", C.POINTER(C.c_int))]


class WakeC(C.Structure):
    _fields_ = [("nw", C.c_int), ("x", C.POINTER(C.c_double)), ("ntri", C.c_int), ("tris", C.POINTER(C.c_int)),
                ("nfil", C.c_int), ("filaments", C.POINTER(C.c_int))]


_PD, _PI = C.POINTER(C.c_double), C.POINTER(C.c_int)


class FreeWakeSetupC(C.Structure):
    """vp_freewake_setup"""
    _fields_ = [("vinf", _PD), ("dt", C.c_double), ("advection", C.c_int), ("delete_beyond", C.c_double),
                ("span", C.c_double), ("keep_rows", C.c_int), ("core_mode", C.c_int), ("core_value", C.c_double),
                ("overlap", C.c_double), ("core_factor", C.c_double), ("core_floor", C.c_double),
                ("opts", C.POINTER(AssemblyOptionsC)), ("field_precision", C.c_int), ("factor", C.c_void_p),
                ("b0", _PD), ("nq", C.c_int), ("C_rowptr", _PI), ("C_col", _PI), ("C_val", _PD), ("nte", C.c_int),
                ("K_rowptr", _PI), ("K_col", _PI), ("K_val", _PD), ("te_x", _PD), ("nteseg", C.c_int),
                ("teseg", _PI), ("edge", _PD), ("nstrip", C.c_int), ("strips", _PI), ("merge", _PI),
                ("nfirst", C.c_int), ("first", _PI), ("nlast", C.c_int), ("last", _PI), ("panel", C.POINTER(WakeC)),
                ("panel_te", _PI), ("tree_theta", C.c_double), ("tree_degree", C.c_int)]


class FreeWakeRecordC(C.Structure):
    """vp_freewake_record"""
    _fields_ = [("n", C.c_int), ("rows", C.c_int), ("particles", C.c_int), ("iters", C.c_int), ("t", C.c_double),
                ("F_pressure", C.c_double * 3), ("M_pressure", C.c_double * 3), ("F_kj", C.c_double * 3),
                ("F_impulse", C.c_double * 3), ("Omega", C.c_double * 3), ("I", C.c_double * 3),
                ("lam", C.c_double), ("umax", C.c_double), ("t_rhs", C.c_double), ("t_vel", C.c_double)]


class TreeOptionsC(C.Structure):
    """vp_tree_options"""
    _fields_ = [("theta", C.c_double), ("degree", C.c_int), ("precision", C.c_int)]


class TreeStatsC(C.Structure):
    """vp_tree_stats"""
    _fields_ = [("far_clusters", C.c_longlong), ("far_pairs", C.c_longlong), ("near_pairs", C.c_longlong),
                ("proxy_evals", C.c_longlong), ("near_evals", C.c_longlong), ("interp_evals", C.c_longlong),
                ("build_ms", C.c_double), ("device_ms", C.c_double), ("sort_ms", C.c_double), ("lists_ms", C.c_double)]


_lib = None

# every symbol include/vortexhip.h declares
EXPORTS = ["vp_create", "vp_destroy", "vp_last_error", "vp_sync", "vp_assemble", "vp_plan_create", "vp_plan_assemble",
           "vp_plan_read", "vp_plan_create_ex", "vp_plan_tiers", "vp_plan_timings", "vp_plan_device_A", "vp_plan_destroy", "vp_factorize",
           "vp_factor_solve", "vp_factor_cond", "vp_factor_info", "vp_factor_destroy", "vp_gemm", "vp_dense_solve",
           "vp_surface_velocity", "vp_body_velocity", "vp_blob_velocities", "vp_plan_adjoint",
           "vp_factor_solve_transposed", "vp_plan_wake_product", "vp_plan_blob_rhs", "vp_gemm_tuning", "vp_math_probe",
           "vp_assemble_mixed", "vp_wake_field_create", "vp_wake_field_eval", "vp_wake_field_info", "vp_wake_field_destroy",
           "vp_plan_update_wake", "vp_assembly_tuning", "vp_factor_mode", "vp_factor_gmres",
           "vp_consistency", "vp_blob_velocities_device", "vp_device_alloc", "vp_device_free", "vp_copy_to_device",
           "vp_copy_from_device", "vp_explicit_create", "vp_explicit_rhs", "vp_explicit_info", "vp_explicit_destroy",
           "vp_diag_lu_probe", "vp_freewake_create", "vp_freewake_step", "vp_freewake_counts", "vp_freewake_read", "vp_freewake_destroy",
           "vp_blob_velocities_tree", "vp_tree_coverage", "vp_factor_inverse"]


def load():
    """Load the built library (raises when it has not been built: the GPU path has no CPU fallback)"""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        lib = C.CDLL(LIBPATH)
        lib.vp_gemm_tuning.restype = C.c_char_p
        lib.vp_gemm_tuning.argtypes = []
        lib.vp_assembly_tuning.restype = C.c_char_p
        lib.vp_assembly_tuning.argtypes = []
        lib.vp_last_error.restype = C.c_char_p
        lib.vp_last_error.argtypes = [C.c_void_p]
        lib.vp_destroy.restype = None
        lib.vp_plan_destroy.restype = None
        lib.vp_wake_field_destroy.restype = None
        lib.vp_explicit_destroy.restype = None
        lib.vp_freewake_destroy.restype = None
        _lib = lib
    return _lib


def ptr(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
