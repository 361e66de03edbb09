"""ctypes view of include/tg.h and include/tg_host.h.

The library is the in-tree discrete_adjoint_amd/libtg.so built by
discrete_adjoint_amd/build.py.  Loading fails loudly when it is missing:
there is no CPU fallback behind this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtg.so")

TG_OK = 0
TG_ERR_ARG = -1
TG_ERR_HIP = -2
TG_ERR_STATE = -3
TG_ERR_NOMEM = -4
TG_ERR_NONFINITE = -5
TG_ERR_EIGEN = -6
TG_ERR_UNSUPPORTED = -7
TG_ERR_NODEVICE = -8
TG_ERR_TOO_LARGE = -9
TG_NOALLOC_MAX_CELLS = 1 << 20

TG_FP64 = 0
TG_MIXED = 1
TG_FP32 = 2
TG_FWD64_ADJ32 = 3
TG_NUM_TILE_VARIANTS = 4
TG_MESH_UNIFORM = 0
TG_MESH_LEAVES = 1

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)


class tg_config(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("checkpoint_every", C.c_int),
        ("max_steps", C.c_int),
        ("device", C.c_int),
        ("tiny", C.c_double),
        ("gamma", C.c_double),
        ("cfl_inner", C.c_double),
        ("cfl_outer", C.c_double),
        ("frict_tiny_early", C.c_double),
        ("frict_tiny_late", C.c_double),
        ("frict_tiny_switch", C.c_int),
        ("tile_x", C.c_int),
        ("tile_y", C.c_int),
        ("use_graph", C.c_int),
        ("store_bytes", C.c_int64),
    ]


class tg_mesh(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("dx", C.c_double),
        ("dy", C.c_double),
        ("x0", C.c_double),
        ("y0", C.c_double),
        ("grav", _dp),
        ("dgrav", _dp),
        ("curv", _dp),
        ("n", C.c_int),
        ("nbr", _i32p),
        ("nproc", _i32p),
        ("gen", _i32p),
        ("cdx", _dp),
        ("xc", _dp),
        ("mat", C.POINTER(C.c_uint8)),
    ]


TG_MAX_MATERIALS = 16


class tg_material(C.Structure):
    _fields_ = [("tan_bed", C.c_double), ("kact_root", C.c_double), ("dkroot_dphi", C.c_double)]


class tg_params(C.Structure):
    _fields_ = [
        ("sin_int", C.c_double),
        ("tan_bed", C.c_double),
        ("kact_c1", C.c_double),
        ("kact_root", C.c_double),
        ("dkroot_dphi", C.c_double),
        ("epsilon", C.c_double),
        ("gravity_scale", C.c_double),
        ("maxtime", C.c_double),
        ("nextoutput", C.c_double),
        ("nextsave", C.c_double),
    ]


class tg_stats(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("time", C.c_double),
        ("mass", C.c_double),
        ("hmax", C.c_double),
        ("step", C.c_int),
    ]


class tg_functional(C.Structure):
    _fields_ = [("terminal_weight", _dp), ("step", C.POINTER(C.c_int))]


class tg_dem(C.Structure):
    _fields_ = [
        ("nrows", C.c_int),
        ("ncols", C.c_int),
        ("xmin", C.c_double),
        ("ymax", C.c_double),
        ("res", C.c_double),
        ("elev", _fp),
    ]


class tg_leaf_table(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("order", C.POINTER(C.c_int32)),
        ("key", C.POINTER(C.c_uint32)),
        ("gen", C.POINTER(C.c_int32)),
        ("nbr", C.POINTER(C.c_int32)),
        ("nkind", C.POINTER(C.c_int32)),
        ("ngen", C.POINTER(C.c_int32)),
        ("dx", _dp),
        ("xc", _dp),
    ]



class tg_error_grid(C.Structure):
    _fields_ = [("n", C.c_int), ("nbr", C.POINTER(C.c_int32)), ("nproc", C.POINTER(C.c_int32)),
                ("gen", C.POINTER(C.c_int32)), ("cdx", _dp), ("xc", _dp), ("grav", _dp),
                ("dgrav", _dp), ("curv", _dp), ("mat", C.POINTER(C.c_uint8)),
                ("father", C.POINTER(C.c_int32)), ("kind", C.POINTER(C.c_int32)),
                ("pts", C.POINTER(C.c_int32)), ("pxy", _dp)]


class tg_noalloc_block(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("statics", C.c_void_p)]


class tg_adapt_params(C.Structure):
    _fields_ = [("max_level", C.c_int), ("num_buffer_layer", C.c_int), ("tiny", C.c_double)]


class tg_adapt_op(C.Structure):
    _fields_ = [("kind", C.c_int32), ("lev", C.c_int32), ("ix", C.c_int32), ("iy", C.c_int32)]


class tg_pile(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("height", "xc", "yc", "major", "minor", "rot_deg",
                                          "cosrot", "sinrot", "vx", "vy")]


class tg_funky_header(C.Structure):
    _fields_ = [("version", C.c_int), ("nnodes", C.c_int), ("nelems", C.c_int),
                ("keyrange", C.c_double * 2), ("xrange", C.c_double * 2),
                ("yrange", C.c_double * 2)]


class tg_funky_elem(C.Structure):
    _fields_ = [("node_key", (C.c_uint32 * 2) * 9), ("neigh_proc", C.c_int32 * 4),
                ("neigh_key", (C.c_uint32 * 2) * 4), ("bc_type", C.c_int32 * 4),
                ("bc_value", ((C.c_float * 2) * 2) * 4), ("elm_loc", C.c_int32 * 2),
                ("opposite_brother", C.c_uint32 * 2), ("material", C.c_int32)]


class tg_run_inputs(C.Structure):
    _fields_ = [("srctype", C.c_int), ("npiles", C.c_int), ("nsources", C.c_int),
                ("length_scale", C.c_double), ("height_scale", C.c_double),
                ("gravity_scale", C.c_double), ("epsilon", C.c_double),
                ("time_scale", C.c_double), ("smallest_axis", C.c_double),
                ("cells_across", C.c_int), ("maxiter", C.c_int), ("maxtime", C.c_double),
                ("timeoutput", C.c_double), ("timesave", C.c_double),
                ("nd_maxtime", C.c_double), ("nd_nextoutput", C.c_double),
                ("nd_nextsave", C.c_double), ("adaptflag", C.c_int), ("vizflag", C.c_int),
                ("order", C.c_int), ("gis_main", C.c_char * 300), ("gis_sub", C.c_char * 100),
                ("gis_mapset", C.c_char * 100), ("gis_map", C.c_char * 100),
                ("extramaps", C.c_int), ("edge_height", C.c_double),
                ("test_height", C.c_double), ("test_x", C.c_double), ("test_y", C.c_double),
                ("nplanes", C.c_int), ("nmat", C.c_int), ("matnames", (C.c_char * 64) * 16),
                ("intfrict", C.c_double), ("bedfrict", C.c_double * 16)]


# name -> (restype, argtypes); every symbol declared in include/*.h
SIGNATURES = {
    "tg_config_default": (None, [C.POINTER(tg_config)]),
    "tg_params_from_angles": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.POINTER(tg_params)]),
    "tg_material_from_angles": (C.c_int, [C.c_double, C.c_double, C.POINTER(tg_material)]),
    "tg_set_materials": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(tg_material)]),
    "tg_set_error_grid": (C.c_int, [C.c_void_p, C.POINTER(tg_error_grid)]),
    "tg_error_indicator": (C.c_int, [C.c_void_p, C.POINTER(tg_params), C.c_int, C.c_double, _dp,
                                     _dp, _dp, _dp, _dp, _dp, _dp]),
    "tg_error_last_ms": (C.c_int, [C.c_void_p, _dp]),
    "tg_create": (C.c_int, [C.POINTER(tg_config), C.POINTER(C.c_void_p)]),
    "tg_destroy": (None, [C.c_void_p]),
    "tg_last_error": (C.c_char_p, [C.c_void_p]),
    "tg_set_mesh": (C.c_int, [C.c_void_p, C.POINTER(tg_mesh)]),
    "tg_set_state": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "tg_get_state": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "tg_set_clock": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "tg_set_ensemble": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(tg_mesh),
                                  C.POINTER(tg_params)]),
    "tg_num_members": (C.c_int, [C.c_void_p]),
    "tg_member_cells": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64)]),
    "tg_set_member_state": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
    "tg_get_member_state": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
    "tg_get_member_dt_record": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_int,
                                          C.POINTER(C.c_int)]),
    "tg_get_member_dphi": (C.c_int, [C.c_void_p, _dp]),
    "tg_forward_step": (C.c_int, [C.c_void_p, C.POINTER(tg_params), C.POINTER(tg_stats)]),
    "tg_forward_run": (C.c_int, [C.c_void_p, C.POINTER(tg_params), C.c_int,
                                 C.POINTER(tg_stats)]),
    "tg_forward_replay": (C.c_int, [C.c_void_p, C.POINTER(tg_params), C.c_int, _dp,
                                    C.POINTER(tg_stats)]),
    "tg_get_dt_record": (C.c_int, [C.c_void_p, _dp, C.c_int, C.POINTER(C.c_int)]),
    "tg_get_trajectory_state": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
    "tg_write_solution_files": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, _u32p, _i32p,
                                          C.c_int, C.c_int]),
    "tg_adjoint_run": (C.c_int, [C.c_void_p, C.POINTER(tg_params), C.POINTER(tg_functional),
                                 _dp, _dp]),
    "tg_set_functional": (C.c_int, [C.c_void_p, C.POINTER(tg_functional)]),
    "tg_set_maxh_region": (C.c_int, [C.c_void_p, C.POINTER(C.c_ubyte)]),
    "tg_get_member_box_record": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_int,
                                           C.POINTER(C.c_int)]),
    "tg_get_maxh": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_int)]),
    "tg_adjoint_step": (C.c_int, [C.c_void_p, C.POINTER(tg_params), C.c_int, C.c_double,
                                  _dp, _dp]),
    "tg_adjoint_step_grad": (C.c_int, [C.c_void_p, C.POINTER(tg_params), C.c_int, C.c_double,
                                       _dp, _dp, _dp]),
    "tg_last_timing": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "tg_last_stored_sweep": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_int)]),
    "tg_time_kernel": (C.c_int, [C.c_void_p, C.POINTER(tg_params), C.c_int, C.c_int, _dp]),
    "tg_num_cells": (C.c_int64, [C.c_void_p]),
    "tg_set_instrumentation": (C.c_int, [C.c_void_p, C.c_int]),
    "tg_get_instrumentation": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_int64),
                                         C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "tg_device_memory": (C.c_int, [C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "tg_tile_variant": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int)]),
    "tg_autotune": (C.c_int, [C.c_void_p, C.POINTER(tg_params), C.c_int, C.POINTER(C.c_int),
                              _dp]),
    "tg_step_noalloc": (C.c_int, [C.POINTER(tg_config), C.POINTER(tg_noalloc_block),
                                  C.POINTER(tg_params), C.c_int, C.c_double, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, _dp]),
    "tg_noalloc_scratch_bytes": (C.c_longlong, []),
    "tg_get_tile_selection": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "tg_set_tile_selection": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    # tg_host.h
    "tg_host_dem_plane_runout": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double,
                                           C.c_double, C.c_double, C.c_double, C.c_double,
                                           _fp]),
    "tg_host_dem_diamond_square": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_uint64,
                                             _fp]),
    "tg_host_static_coefs": (C.c_int, [C.POINTER(tg_dem), C.c_int, C.c_int, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, _dp, _dp, _dp]),
    "tg_host_read_grass": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                     C.c_longlong, C.POINTER(tg_dem), _fp]),
    "tg_host_read_run_inputs": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(tg_pile),
                                          C.POINTER(tg_run_inputs)]),
    "tg_host_read_funky": (C.c_int, [C.c_char_p, C.c_double, C.POINTER(tg_funky_header),
                                     C.c_int, C.POINTER(C.c_uint32), _dp, C.c_int,
                                     C.POINTER(tg_funky_elem)]),
    "tg_host_leaf_static_coefs": (C.c_int, [C.POINTER(tg_dem), C.c_int, _dp, C.c_double,
                                            C.c_double, _dp, _dp]),
    "tg_host_adapt": (C.c_int, [C.c_int, C.c_int, C.POINTER(tg_adapt_params), C.c_int, _i32p,
                                _i32p, _i32p, _dp, _dp, _dp, _dp, C.c_int, C.POINTER(C.c_int),
                                _i32p, _i32p, _i32p, _dp, _dp, _dp, C.c_int,
                                C.POINTER(C.c_int), C.POINTER(tg_adapt_op)]),
    "tg_host_adapt_forward": (C.c_int, [C.c_int, C.POINTER(tg_adapt_op), C.c_int, C.c_int,
                                        _i32p, _i32p, _i32p, _dp, C.c_int, _i32p, _i32p, _i32p,
                                        _dp]),
    "tg_host_adapt_adjoint": (C.c_int, [C.c_int, C.POINTER(tg_adapt_op), C.c_int, _i32p, _i32p,
                                        _i32p, _dp, C.c_int, _i32p, _i32p, _i32p, _dp]),
    "tg_host_piles": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                _dp, _dp]),
    "tg_host_box_weights": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_double, _dp]),
    "tg_host_leaf_d_gravity": (C.c_int, [C.c_int, _i32p, _dp, _dp, _dp]),
    "tg_host_write_solution": (C.c_int, [C.c_char_p, C.c_int64, _u32p, _i32p, _dp, _dp, _dp]),
    "tg_host_read_solution": (C.c_int, [C.c_char_p, C.c_int64, _u32p, _dp, _dp, _dp,
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "tg_host_hsfc2d": (C.c_int, [C.c_int, _dp, _u32p]),
    "tg_host_hash_bucket": (C.c_int, [C.c_int, _u32p, C.c_double, C.c_double, C.c_int,
                                      _i32p]),
    "tg_host_traversal": (C.c_int, [C.c_int, _u32p, C.c_double, C.c_double, C.c_int,
                                    _i32p]),
    "tg_host_error_stencil": (C.c_int, [C.c_int, _i32p, _i32p, _dp, _dp, _i32p, _i32p, C.c_int,
                                        _i32p, _i32p, _dp]),
    "tg_host_flatten": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_int, _i32p, _i32p, _i32p, C.c_int,
                                  C.POINTER(tg_leaf_table)]),
}

_LIB = None


def load():
    """Returns the loaded library; raises if libtg.so is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m discrete_adjoint_amd.build`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def dptr(a):
    """Pointer to a C-contiguous float64 numpy array."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "need contiguous float64"
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], "need contiguous int32"
    return a.ctypes.data_as(_i32p)


def uptr(a):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"], "need contiguous uint32"
    return a.ctypes.data_as(_u32p)


def fptr(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], "need contiguous float32"
    return a.ctypes.data_as(_fp)
