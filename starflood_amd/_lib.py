"""ctypes binding of libstarflood_hip.so (include/starflood.h,
include/starflood_ext.h).

The library is loaded from the package's lib/ directory, where
starflood_amd.build places it.  Loading fails loudly when it is missing:
there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libstarflood_hip.so"

STARFLOOD_SUCCESS = 0
STARFLOOD_FAILURE = -1

# enum sim_conf (reference src/simulation.h:17-51)
SIM_POS, SIM_VEL, SIM_ACC = 0, 3, 6
SIM_MAS, SIM_POT, SIM_KEN, SIM_PEN = 9, 10, 11, 12
SIM_RAD, SIM_RHO, SIM_PRS = 13, 14, 15

SF_PRECISION_EXACT, SF_PRECISION_FAST = 0, 1
SF_ACC_REFERENCE, SF_ACC_PHYSICAL = 0, 1
SF_TREE_REFERENCE, SF_TREE_LBVH = 0, 1
SF_SPH_REFERENCE, SF_SPH_CUBIC_IDEAL = 0, 1
SF_SPH_SEARCH_ALL, SF_SPH_SEARCH_GRID = 0, 1
SF_MEM_MIRROR, SF_MEM_RESIDENT = 0, 1
SF_SYNC_TO_DEVICE, SF_SYNC_TO_HOST = 0, 1
SF_EXT_U, SF_EXT_DUDT, SF_EXT_X64 = 0, 1, 2
SF_NOALLOC_MAX_N = 16384   # include/starflood_ext.h


class NodeT(C.Structure):
    """node_t, reference src/tree.h:16-25 (80 bytes)."""
    _fields_ = [("parent", C.c_uint32), ("bodies", C.c_uint32), ("branch", C.c_uint32 * 8),
                ("bounds", C.c_float * 6), ("params", C.c_float * 4)]


class TreeT(C.Structure):
    """tree_t, reference src/tree.h:28-35."""
    _fields_ = [("max_nodes", C.c_uint32), ("num_nodes", C.c_uint32), ("bounds", C.c_float * 6),
                ("node", C.POINTER(NodeT))]


class SimT(C.Structure):
    """sim_t, reference src/simulation.h:53-67 (ENABLE_TREE defined)."""
    _fields_ = [("step_number", C.c_uint), ("N", C.c_uint), ("mem", C.c_void_p), ("tree", TreeT)]


class Params(C.Structure):
    """sf_params_t."""
    _fields_ = [("epsilon", C.c_double), ("G", C.c_double), ("dt", C.c_double), ("theta", C.c_double),
                ("tree_nodes_max", C.c_uint), ("tree_depth_max", C.c_uint), ("leaf_size", C.c_uint),
                ("enable_tree", C.c_int), ("enable_sph", C.c_int), ("kahan_solver", C.c_int),
                ("kahan_energy", C.c_int), ("precision", C.c_int), ("acc_mode", C.c_int),
                ("tree_mode", C.c_int), ("sph_model", C.c_int), ("mem_mode", C.c_int),
                ("sph_gamma", C.c_double), ("logging", C.c_int), ("verbose", C.c_int),
                ("out_dir", C.c_char * 256), ("quadrupole", C.c_int), ("sph_search", C.c_int),
                ("mac_alpha", C.c_double)]


class VisT(C.Structure):
    """vis_t, reference src/visualization.h:18-27."""
    _fields_ = [("sizex", C.c_uint), ("sizey", C.c_uint), ("mem", C.c_void_p), ("atomic_buffer", C.c_void_p),
                ("render_buffer", C.c_void_p), ("binary_buffer", C.c_void_p)]


IMAGE_FILETYPE_PFM, IMAGE_FILETYPE_PPM, IMAGE_FILETYPE_PNG = 0, 1, 2


class RunConfig(C.Structure):
    """sf_run_config_t: the config.h settings of the reference program's run
    loop (src/main.c)."""
    _fields_ = [("num_bodies", C.c_uint), ("num_timesteps", C.c_uint), ("output_interval", C.c_uint),
                ("enable_sim", C.c_int), ("enable_vis", C.c_int), ("sim_filename", C.c_char_p),
                ("vis_filename", C.c_char_p), ("vis_file_type", C.c_int), ("vis_sizex", C.c_uint),
                ("vis_sizey", C.c_uint), ("init_cond_file", C.c_char_p), ("verbose", C.c_int)]


class Stats(C.Structure):
    """sf_stats_t."""
    _fields_ = [("T", C.c_double), ("V", C.c_double), ("E", C.c_double), ("W", C.c_double), ("U", C.c_double),
                ("P", C.c_double * 3), ("L", C.c_double * 3), ("K", C.c_double)]


# (name, restype, argtypes) for every exported entry point
_SIG = [
    ("vis_init", C.c_int, [C.POINTER(VisT), C.c_uint, C.c_uint]),
    ("vis_free", C.c_int, [C.POINTER(VisT)]),
    ("vis_save", C.c_int, [C.POINTER(VisT), C.c_char_p, C.c_int]),
    ("vis_draw", C.c_int, [C.POINTER(VisT), C.POINTER(SimT)]),
    ("sim_find", C.POINTER(C.c_float), [C.POINTER(SimT), C.c_int]),
    ("sim_init", C.c_int, [C.POINTER(SimT), C.c_uint]),
    ("sim_free", C.c_int, [C.POINTER(SimT)]),
    ("sim_read_raw", C.c_int, [C.POINTER(SimT), C.c_char_p]),
    ("sim_save_raw", C.c_int, [C.POINTER(SimT), C.c_char_p]),
    ("sim_solv", C.c_int, [C.POINTER(SimT)]),
    ("sim_step", C.c_int, [C.POINTER(SimT)]),
    ("solve_gravity_part_tree", C.c_int, [C.POINTER(SimT)]),
    ("solve_gravity_part_part", C.c_int, [C.POINTER(SimT)]),
    ("solve_sph", C.c_int, [C.POINTER(SimT)]),
    ("tree_init", C.c_int, [C.POINTER(TreeT), C.c_uint]),
    ("tree_free", C.c_int, [C.POINTER(TreeT)]),
    ("tree_build", C.c_int, [C.POINTER(TreeT), C.c_uint, C.c_void_p, C.c_void_p]),
    ("initcond_generate", C.c_int, [C.POINTER(SimT)]),
    ("sf_params_default", None, [C.POINTER(Params)]),
    ("sf_set_params", C.c_int, [C.POINTER(Params)]),
    ("sf_get_params", C.c_int, [C.POINTER(Params)]),
    ("sf_last_error", C.c_char_p, []),
    ("sf_version", C.c_char_p, []),
    ("sf_sim_dof", C.c_uint, []),
    ("sf_sync", C.c_int, [C.POINTER(SimT), C.c_int]),
    ("sf_device_find", C.c_void_p, [C.POINTER(SimT), C.c_int]),
    ("sf_device_ext", C.c_void_p, [C.POINTER(SimT), C.c_int]),
    ("sf_ext_copy", C.c_int, [C.POINTER(SimT), C.c_int, C.c_void_p, C.c_int]),
    ("sf_get_stats", C.c_int, [C.POINTER(SimT), C.POINTER(Stats)]),
    ("sf_ic_uniform_ball", C.c_int, [C.POINTER(SimT), C.c_uint32, C.c_double, C.c_double, C.c_double]),
    ("sf_ic_plummer", C.c_int, [C.POINTER(SimT), C.c_uint32, C.c_uint, C.c_uint, C.c_double, C.c_double,
                                C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("sf_lbvh_export", C.c_int, [C.POINTER(TreeT), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p]),
    ("sf_prof_enable", None, [C.c_int]),
    ("sf_prof_reset", None, []),
    ("sf_prof_query", C.c_int, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_ulong)]),
    ("sf_pp_direct_f64", C.c_int, [C.POINTER(SimT), C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p]),
    ("sf_sph_direct_f64", C.c_int, [C.POINTER(SimT), C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("sf_adjoint_forward", C.c_int, [C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint, C.c_uint,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("sf_adjoint_backward", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("sf_tape_free", None, [C.c_void_p]),
    ("sf_force_dev", C.c_int, [C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("sf_force_vjp_dev", C.c_int, [C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("sf_force_dev_noalloc", C.c_int, [C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("sf_run", C.c_int, [C.POINTER(RunConfig)]),
]

_lib = None


def load() -> C.CDLL:
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m starflood_amd.build` "
                               "(there is no fallback implementation)")
        # local binding: the reference's symbol names must not leak into
        # other libraries loaded later in the process
        lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | C.RTLD_LOCAL)
        for name, res, args in _SIG:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def exported_names() -> list[str]:
    return [name for name, _, _ in _SIG]


def last_error() -> str:
    return (load().sf_last_error() or b"").decode()


def check(rc: int, what: str) -> None:
    if rc != STARFLOOD_SUCCESS:
        raise RuntimeError(f"{what} failed: {last_error()}")


def default_params() -> Params:
    p = Params()
    load().sf_params_default(C.byref(p))
    return p


def set_params(**kw) -> Params:
    """Defaults (the reference's config.h values) overridden by kw."""
    p = default_params()
    for k, v in kw.items():
        if k == "out_dir":
            v = v.encode() if isinstance(v, str) else v
        setattr(p, k, v)
    check(load().sf_set_params(C.byref(p)), "sf_set_params")
    return p
