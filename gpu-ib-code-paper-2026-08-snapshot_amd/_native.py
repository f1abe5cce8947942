"""ctypes binding of ``libpyib_hip.so`` (C ABI declared in ``include/pyib_hip.h``).

The library is built in-tree by ``build.py``.  Loading fails loudly when it is
missing: there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpyib_hip.so")

F64, F32 = 0, 1
MAX_PARTIALS = 8192
NSCAL = 8
S_ETA, S_ETA_NEW, S_DEN, S_ALPHA, S_BETA, S_RR, S_ERR, S_L2I = range(8)

vp = C.c_void_p
i32p = C.POINTER(C.c_int32)


class Mesh(C.Structure):
    _fields_ = [
        ("dtype", C.c_int32), ("reserved", C.c_int32),
        ("ncell", C.c_int64), ("nface", C.c_int64), ("nedge", C.c_int64), ("nghost", C.c_int64),
        ("curl_ptr", vp), ("curl_idx", vp), ("rot_ptr", vp), ("rot_idx", vp),
        ("cf_ptr", vp), ("cf_mid", vp), ("cf_face", vp), ("cf_other", vp),
        ("f2c", vp), ("rf1", vp), ("rf2", vp), ("aol", vp), ("vol_ci", vp),
        ("gh_face", vp), ("gh_cell", vp), ("gh_kind", vp), ("gh_nf", vp), ("gh_af2", vp),
        ("ninlet", C.c_int64), ("inlet_edge", vp), ("inlet_val", vp), ("dir_mask", vp),
    ]


MAX_LEVELS = 8
IB_NOALLOC_MAX_MARKERS = 16384     # PYIB_IB_NOALLOC_MAX_MARKERS of include/pyib_hip.h


class Level(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("reserved", C.c_int32),
                ("cell_off", C.c_int64), ("e_off", C.c_int64 * 3), ("h", C.c_double),
                ("hole_lo", C.c_int32 * 3), ("hole_hi", C.c_int32 * 3)]


class Grid(C.Structure):
    _fields_ = [
        ("dtype", C.c_int32), ("nlevels", C.c_int32), ("ncell", C.c_int64), ("nedge", C.c_int64),
        ("lev", Level * MAX_LEVELS), ("cell_reg", vp), ("cell_reg_g1", vp), ("edge_reg", vp), ("dir", vp),
        ("cell_g1v", vp),
        ("g1_nslice", C.c_int64), ("g1_row", vp), ("g1_width", vp), ("g1_sptr", vp), ("g1_col", vp), ("g1_coef", vp),
        ("v_nslice", C.c_int64), ("v_row", vp), ("v_width", vp), ("v_sptr", vp), ("v_col", vp), ("v_coef", vp),
        ("v_rw", vp),
        ("g3_nslice", C.c_int64), ("g3_row", vp), ("g3_width", vp), ("g3_sptr", vp), ("g3_col", vp), ("g3_coef", vp),
        ("edge_cls", vp), ("cls_prec", vp), ("prec_level", C.c_double * MAX_LEVELS),
        ("prec_level_g3", C.c_double * MAX_LEVELS), ("g3_mask", vp), ("g3_rprec", vp),
    ]


class Work(C.Structure):
    _fields_ = [("face", vp), ("cell_a", vp), ("cell_b", vp), ("partials", vp)]


class PGraph(C.Structure):
    _fields_ = [("ncell", C.c_int64), ("nface", C.c_int64), ("fc", vp), ("pos", vp), ("grad", vp), ("area", vp),
                ("ptr", vp), ("nbr", vp), ("sface", vp), ("gauge", vp), ("w", vp), ("jmp", vp), ("dinv", vp),
                ("x", vp), ("r", vp), ("z", vp), ("p", vp), ("q", vp), ("scal", vp), ("partials", vp)]


class CG(C.Structure):
    _fields_ = [("x", vp), ("r", vp), ("p", vp), ("w", vp), ("prec", vp), ("scal", vp)]


_SIGS = {
    "pyib_last_error": (C.c_char_p, []),
    "pyib_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "pyib_curl": (C.c_int, [C.POINTER(Mesh), vp, vp, vp]),
    "pyib_rot": (C.c_int, [C.POINTER(Mesh), vp, vp, vp]),
    "pyib_interp_f2c": (C.c_int, [C.POINTER(Mesh), vp, vp, vp]),
    "pyib_integr_c2f": (C.c_int, [C.POINTER(Mesh), vp, vp, vp]),
    "pyib_cdiv": (C.c_int, [C.POINTER(Mesh), vp, vp, vp]),
    "pyib_flow_mat": (C.c_int, [C.POINTER(Mesh), C.POINTER(Work), C.c_double, C.c_double, C.c_double,
                                vp, vp, C.c_int, vp, vp]),
    "pyib_calc_vel": (C.c_int, [C.POINTER(Mesh), vp, vp, vp, vp, vp, vp]),
    "pyib_momentum_residual": (C.c_int, [C.POINTER(Mesh), vp, vp, C.c_double, vp, C.c_double, C.c_double,
                                         C.c_double, vp, C.c_double, vp, vp]),
    "pyib_curl_rhs": (C.c_int, [C.POINTER(Mesh), C.POINTER(Work), vp, vp, vp]),
    "pyib_cg_begin": (C.c_int, [C.POINTER(Mesh), C.POINTER(Work), C.POINTER(CG), vp, vp, C.c_double,
                                C.c_double, vp]),
    "pyib_cg_iterate": (C.c_int, [C.POINTER(Mesh), C.POINTER(Work), C.POINTER(CG), C.c_double, C.c_double,
                                  C.c_double, C.c_int32, C.c_int32, vp]),
    "pyib_cg_direction": (C.c_int, [C.POINTER(Mesh), C.POINTER(Work), C.POINTER(CG), vp]),
    "pyib_cg_true_residual": (C.c_int, [C.POINTER(Mesh), C.POINTER(Work), C.POINTER(CG), vp, vp, vp, vp,
                                        C.c_double, C.c_double, C.c_double, vp, vp]),
    "pyib_grid_flow_mat": (C.c_int, [C.POINTER(Grid), C.POINTER(Work), C.c_double, C.c_double, C.c_double,
                                     vp, vp, vp, vp]),
    "pyib_grid_cg_begin": (C.c_int, [C.POINTER(Grid), C.POINTER(Work), C.POINTER(CG), vp, vp, C.c_double,
                                     C.c_double, vp]),
    "pyib_grid_cg_iterate": (C.c_int, [C.POINTER(Grid), C.POINTER(Work), C.POINTER(CG), C.c_double, C.c_double,
                                       C.c_double, C.c_int32, C.c_int32, vp]),
    "pyib_grid_cg_direction": (C.c_int, [C.POINTER(Grid), C.POINTER(CG), vp]),
    "pyib_grid_cg_true_residual": (C.c_int, [C.POINTER(Grid), C.POINTER(Work), C.POINTER(CG), vp, vp, vp, vp,
                                             C.c_double, C.c_double, C.c_double, vp, vp]),
    "pyib_permute": (C.c_int, [C.c_int, vp, C.c_int, vp, C.c_int64, vp, vp, vp]),
    "pyib_extrapolate": (C.c_int, [C.c_int64, vp, vp, C.c_double, C.c_double, vp, vp]),
    "pyib_convert": (C.c_int, [C.c_int, vp, C.c_int, vp, C.c_int64, vp]),
    "pyib_cg_replace_residual": (C.c_int, [vp, vp, vp]),
    "pyib_add_into_f64": (C.c_int, [C.c_int, vp, vp, C.c_int64, vp]),
    "pyib_dot": (C.c_int, [C.c_int, vp, vp, vp, C.c_int64, vp, vp, vp]),
    "pyib_axpby": (C.c_int, [C.c_int, C.c_double, vp, C.c_double, vp, vp, C.c_int64, vp]),
    "pyib_ib_stencil": (C.c_int, [C.c_int64, vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_int32, C.c_int32, C.c_int32, vp, vp, vp]),
    "pyib_ib_interp": (C.c_int, [C.c_int, C.c_int64, vp, vp, C.c_double, vp, vp, vp]),
    "pyib_ib_sort_temp_bytes": (C.c_int, [C.c_int64, C.POINTER(C.c_int64)]),
    "pyib_ib_sort": (C.c_int, [C.c_int64, vp, vp, vp, vp, vp, vp, C.c_int64, vp]),
    "pyib_ib_sort_noalloc": (C.c_int, [C.c_int64, vp, vp, vp, vp]),
    "pyib_ib_spread": (C.c_int, [C.c_int, C.c_int64, vp, vp, vp, vp, vp, C.c_double, C.c_int, vp, vp]),
    "pyib_ib_coupling": (C.c_int, [C.c_int64, vp, vp, vp, vp, vp, C.c_double, vp, vp]),
    "pyib_ib_force": (C.c_int, [C.c_int64, vp, vp, C.c_double, vp, vp]),
    "pyib_interior_momenta": (C.c_int, [C.c_int, C.c_int64, vp, vp, vp, vp, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_int, vp, vp, vp, vp]),
    "pyib_fluid_load": (C.c_int, [C.c_int, C.c_int64, vp, vp, vp, vp, vp, vp, vp]),
    "pyib_diag_max": (C.c_int, [C.POINTER(Mesh), vp, vp, vp, vp, vp]),
    "pyib_calc_vel_adjoint": (C.c_int, [C.POINTER(Mesh), vp, vp, vp, vp, vp, vp]),
    "pyib_curl_rhs_adjoint": (C.c_int, [C.POINTER(Mesh), C.POINTER(Work), vp, vp, vp]),
    "pyib_momentum_residual_adjoint": (C.c_int, [C.POINTER(Mesh), vp, vp, vp, C.c_double, C.c_int, C.c_double,
                                                 C.c_double, C.c_double, vp, vp, vp, vp]),
    "pyib_ib_dweights": (C.c_int, [C.c_int64, vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_int32, C.c_int32, C.c_int32, vp, vp]),
    "pyib_ib_marker_grad": (C.c_int, [C.c_int64, vp, vp, vp, vp, vp, C.c_double, vp, vp]),
    "pyib_ib_gather": (C.c_int, [C.c_int64, vp, vp, vp, vp, vp, vp]),
    "pyib_fluid_load_adjoint": (C.c_int, [C.c_int64, vp, vp, C.c_double, C.c_double, C.c_double, vp, vp]),
    "pyib_geometry": (C.c_int, [C.c_int64, vp, C.c_int64, vp, C.c_int64, vp, vp, C.c_int64, vp, C.c_int64,
                                vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "pyib_tables": (C.c_int, [C.c_int64, C.c_int64, C.c_int64, C.c_int64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                              vp, vp]),
    "pyib_pressure_graph": (C.c_int, [C.c_int64, C.c_int64, vp, vp, vp, vp]),
    "pyib_pressure_begin": (C.c_int, [C.POINTER(PGraph), vp]),
    "pyib_pressure_iterate": (C.c_int, [C.POINTER(PGraph), C.c_int32, vp]),
    "pyib_pressure_finish": (C.c_int, [C.POINTER(PGraph), vp, vp]),
    "pyib_mesh_build": (C.c_int, [C.c_int32, vp, vp, vp, vp, C.POINTER(vp), vp]),
    "pyib_mesh_fetch": (C.c_int, [vp] + [vp] * 15),
    "pyib_mesh_free": (C.c_int, [vp]),
    "pyib_layout_build": (C.c_int, [C.c_int32, vp, vp, vp, C.c_int64, C.c_int64, C.c_int64, vp, C.c_int64, vp, vp,
                                    C.c_int64, C.c_int64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                    vp, C.c_int32, C.c_double, C.c_int32, C.POINTER(vp), vp]),
    "pyib_layout_fetch": (C.c_int, [vp] + [vp] * 20),
    "pyib_layout_free": (C.c_int, [vp]),
    "pyib_interior_momentum_adjoint": (C.c_int, [C.c_int64, vp, vp, vp, vp, C.c_double, C.c_double, C.c_double,
                                                 C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                                 C.c_double, vp, vp, vp, vp]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def load(path: str = None):
    """Load the HIP library (raises if it has not been built).  PYIB_HIP_LIBRARY overrides
    the in-tree path (a library built elsewhere from these sources)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("PYIB_HIP_LIBRARY") or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library {path} is missing; run __graft_entry__.build() (no CPU fallback exists)")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def call(name, *args):
    """Call ``name`` and raise RuntimeError with the library's message on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.pyib_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed ({rc}): {msg}")
    return rc
