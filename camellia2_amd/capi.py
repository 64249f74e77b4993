"""ctypes mirror of include/dpg_capi.h and loader of libcamellia2_dpg.so.

The structs here are byte-for-byte the C structs of the ABI; the loader
raises if the shared library is missing (there is no Python fallback).
"""
import ctypes as ct
import os

import numpy as np

MAX_VARS = 16
MAX_BTERMS = 64
MAX_IPTERMS = 32
MAX_LT_SUMMANDS = 8
MAX_RHS_TERMS = 8
MAX_LOAD_MODES = 32
MAX_COEFS = 8
NOALLOC_SCRATCH_BYTES = 256 << 20  # DPG_NOALLOC_SCRATCH_BYTES

QUAD, HEX, TRI = 0, 1, 2
FIELD, FIELD_VECTOR, TRACE, FLUX = 0, 1, 2, 3
HGRAD, HDIV = 0, 1
OP_VALUE, OP_DX, OP_DY, OP_DZ, OP_X, OP_Y, OP_Z, OP_DIV, OP_DOT_NORMAL = range(9)
FN_VALUE, FN_SQRT, FN_INV = 0, 1, 2
PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_FP16X3, PREC_INT8SPLIT = 0, 1, 2, 3, 4


class Coef(ct.Structure):
    _fields_ = [("scale", ct.c_double), ("coef_slot", ct.c_int32), ("coef_fn", ct.c_int32),
                ("normal_comp", ct.c_int32), ("pad", ct.c_int32)]


class BTerm(ct.Structure):
    _fields_ = [("trial_var", ct.c_int32), ("trial_comp", ct.c_int32), ("test_var", ct.c_int32),
                ("test_op", ct.c_int32), ("coef", Coef)]


class LtSummand(ct.Structure):
    _fields_ = [("test_var", ct.c_int32), ("test_op", ct.c_int32), ("coef", Coef)]


class IpTerm(ct.Structure):
    _fields_ = [("n", ct.c_int32), ("pad", ct.c_int32), ("s", LtSummand * MAX_LT_SUMMANDS)]


class RhsTerm(ct.Structure):
    _fields_ = [("test_var", ct.c_int32), ("test_op", ct.c_int32), ("load_comp", ct.c_int32),
                ("pad", ct.c_int32), ("scale", ct.c_double)]


class FormDesc(ct.Structure):
    _fields_ = [("spatial_dim", ct.c_int32), ("n_trial", ct.c_int32), ("n_test", ct.c_int32),
                ("n_coefs", ct.c_int32), ("trial_kind", ct.c_int32 * MAX_VARS),
                ("test_space", ct.c_int32 * MAX_VARS), ("n_bterms", ct.c_int32),
                ("n_ipterms", ct.c_int32), ("n_rhs", ct.c_int32), ("pad", ct.c_int32),
                ("bterms", BTerm * MAX_BTERMS), ("ipterms", IpTerm * MAX_IPTERMS),
                ("rhs", RhsTerm * MAX_RHS_TERMS)]


class LoadDesc(ct.Structure):
    _fields_ = [("ncomp", ct.c_int32), ("nmodes", ct.c_int32 * 3), ("constant", ct.c_double * 3),
                ("amp", (ct.c_double * MAX_LOAD_MODES) * 3),
                ("k", ((ct.c_int32 * 3) * MAX_LOAD_MODES) * 3)]


class BucketDesc(ct.Structure):
    _fields_ = [("topology", ct.c_int32), ("h1_order", ct.c_int32), ("test_degree", ct.c_int32),
                ("pad", ct.c_int32)]


class BucketInfo(ct.Structure):
    _fields_ = [("N", ct.c_int32), ("M", ct.c_int32), ("F", ct.c_int32), ("T", ct.c_int32),
                ("Q", ct.c_int32), ("Qs", ct.c_int32), ("QG", ct.c_int32), ("nsides", ct.c_int32),
                ("nverts", ct.c_int32), ("pad", ct.c_int32), ("flops_per_cell", ct.c_double)]


class Options(ct.Structure):
    _fields_ = [("precision", ct.c_int32), ("refine", ct.c_int32), ("max_batch_cells", ct.c_int32),
                ("profile", ct.c_int32), ("workspace_bytes", ct.c_double)]


STAGES = ["point", "operand", "gram", "bmat", "scale", "panel", "update", "extract", "adjoint", "solve"]
NSTAGES = len(STAGES)


class Timings(ct.Structure):
    _fields_ = [("ms", ct.c_double * NSTAGES), ("flops", ct.c_double * NSTAGES), ("bytes", ct.c_double * NSTAGES),
                ("launches", ct.c_int64 * NSTAGES), ("cells", ct.c_int64)]


class Batch(ct.Structure):
    _fields_ = [("ncells", ct.c_int32), ("pad", ct.c_int32), ("nodes", ct.c_void_p),
                ("parities", ct.c_void_p), ("coefs", ct.c_void_p), ("load", LoadDesc)]


class Outputs(ct.Structure):
    _fields_ = [("K", ct.c_void_p), ("rhs", ct.c_void_p), ("G", ct.c_void_p), ("schur", ct.c_void_p),
                ("schur_rhs", ct.c_void_p), ("info", ct.c_void_p), ("factor", ct.c_void_p),
                ("factor_scale", ct.c_void_p)]


class Tiles(ct.Structure):
    _fields_ = [("gram", ct.c_int32), ("bmat", ct.c_int32), ("update", ct.c_int32), ("sumfact", ct.c_int32)]


class WorkItem(ct.Structure):
    _fields_ = [("bucket", ct.c_int32), ("pad", ct.c_int32), ("batch", Batch), ("out", Outputs)]


LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libcamellia2_dpg.so")

# every entry point declared in include/dpg_capi.h
EXPORTS = [
    "dpg_version", "dpg_last_error", "dpg_form_poisson_uw", "dpg_form_cdr_uw",
    "dpg_form_stokes_vgp_uw", "dpg_form_convection_diffusion_uw", "dpg_form_graph_norm", "dpg_mesh_quad", "dpg_mesh_hex",
    "dpg_mesh_quad_device", "dpg_mesh_hex_device",
    "dpg_load_sine_modes", "dpg_rng_uniform", "dpg_rng_normal", "dpg_bucket_sizes",
    "dpg_plan_create", "dpg_plan_destroy", "dpg_plan_clone_workspace", "dpg_plan_workspace", "dpg_plan_timings",
    "dpg_assemble", "dpg_assemble_noalloc",
    "dpg_energy_error", "dpg_energy_error_factored", "dpg_adjoint", "dpg_mesh_quad_graded", "dpg_plan_get_tiles", "dpg_plan_set_tiles",
    "dpg_plan_autotune", "dpg_graph_create", "dpg_graph_launch", "dpg_graph_workspace", "dpg_graph_destroy",
    "dpg_skeleton_numbering", "dpg_skeleton_nodes", "dpg_skeleton_create", "dpg_skeleton_create_ragged",
    "dpg_skeleton_info",
    "dpg_skeleton_pattern", "dpg_skeleton_assemble", "dpg_skeleton_solve", "dpg_skeleton_entities",
    "dpg_skeleton_set_blocks", "dpg_mark_cells", "dpg_skeleton_destroy",
    "dpg_coarse_create", "dpg_coarse_info", "dpg_coarse_destroy", "dpg_skeleton_solve_twolevel",
    "dpg_hp_numbering", "dpg_hp_system_create", "dpg_hp_system_info", "dpg_hp_system_dofs", "dpg_hp_system_assemble",
    "dpg_hp_system_local", "dpg_hp_system_destroy",
    "dpg_recover_fields", "dpg_riesz_representation", "dpg_side_parities", "dpg_sort_cells", "dpg_gather_cells",
]

_lib = None


def lib():
    """The loaded C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ct.CDLL(LIB_PATH)
        P = ct.POINTER
        L.dpg_version.restype = ct.c_char_p
        L.dpg_last_error.restype = ct.c_char_p
        for name in ("dpg_form_poisson_uw", "dpg_form_cdr_uw", "dpg_form_convection_diffusion_uw"):
            getattr(L, name).argtypes = [ct.c_int, P(FormDesc)]
        L.dpg_form_stokes_vgp_uw.argtypes = [ct.c_int, ct.c_double, P(FormDesc)]
        L.dpg_form_graph_norm.argtypes = [P(FormDesc)]
        L.dpg_mesh_quad.argtypes = [ct.c_int, ct.c_int, ct.c_double, ct.c_uint64, ct.c_void_p, ct.c_void_p]
        L.dpg_mesh_hex.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_double, ct.c_uint64, ct.c_void_p,
                                   ct.c_void_p]
        L.dpg_mesh_quad_device.argtypes = [ct.c_int, ct.c_int, ct.c_double, ct.c_uint64, ct.c_void_p, ct.c_void_p,
                                           ct.c_void_p]
        L.dpg_mesh_hex_device.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_double, ct.c_uint64, ct.c_void_p,
                                          ct.c_void_p, ct.c_void_p]
        L.dpg_mesh_quad_graded.argtypes = [ct.c_int, ct.c_int, ct.c_double, ct.c_uint64, ct.c_void_p,
                                           ct.c_void_p, ct.c_void_p]
        L.dpg_load_sine_modes.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_uint64, P(LoadDesc)]
        L.dpg_rng_uniform.argtypes = [ct.c_uint64, ct.c_double, ct.c_double, ct.c_void_p, ct.c_size_t]
        L.dpg_rng_normal.argtypes = [ct.c_uint64, ct.c_void_p, ct.c_size_t]
        L.dpg_bucket_sizes.argtypes = [P(FormDesc), P(BucketDesc), P(BucketInfo)]
        L.dpg_plan_create.argtypes = [P(FormDesc), P(BucketDesc), ct.c_int, P(Options), P(ct.c_void_p)]
        L.dpg_plan_destroy.argtypes = [ct.c_void_p]
        L.dpg_plan_clone_workspace.argtypes = [ct.c_void_p, ct.c_double, P(ct.c_void_p)]
        L.dpg_plan_workspace.argtypes = [ct.c_void_p, P(ct.c_size_t)]
        L.dpg_plan_timings.argtypes = [ct.c_void_p, P(Timings), ct.c_int]
        L.dpg_assemble.argtypes = [ct.c_void_p, ct.c_int, P(Batch), P(Outputs), ct.c_void_p]
        L.dpg_assemble_noalloc.argtypes = [ct.c_void_p, ct.c_int, P(Batch), P(Outputs), ct.c_void_p]
        L.dpg_energy_error.argtypes = [ct.c_void_p, ct.c_int, P(Batch), ct.c_void_p, ct.c_void_p,
                                       ct.c_void_p]
        L.dpg_energy_error_factored.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p,
                                                ct.c_void_p, ct.c_void_p]
        L.dpg_adjoint.argtypes = [ct.c_void_p, ct.c_int, P(Batch), ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                  ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.dpg_plan_get_tiles.argtypes = [ct.c_void_p, ct.c_int, P(Tiles)]
        L.dpg_plan_set_tiles.argtypes = [ct.c_void_p, ct.c_int, P(Tiles)]
        L.dpg_plan_autotune.argtypes = [ct.c_void_p, ct.c_int, P(Batch), ct.c_int, ct.c_void_p]
        L.dpg_graph_create.argtypes = [ct.c_void_p, P(WorkItem), ct.c_int, ct.c_int, ct.c_void_p,
                                       P(ct.c_void_p)]
        L.dpg_graph_launch.argtypes = [ct.c_void_p, ct.c_void_p]
        L.dpg_graph_workspace.argtypes = [ct.c_void_p, P(ct.c_size_t)]
        L.dpg_graph_destroy.argtypes = [ct.c_void_p]
        L.dpg_skeleton_numbering.argtypes = [P(FormDesc), P(BucketDesc), ct.c_int, P(ct.c_int32), P(ct.c_int32)]
        L.dpg_skeleton_entities.argtypes = [P(FormDesc), P(BucketDesc), ct.c_int, P(ct.c_int32), P(ct.c_int32)]
        L.dpg_skeleton_set_blocks.argtypes = [ct.c_void_p, P(ct.c_int32), ct.c_int, ct.c_void_p]
        L.dpg_skeleton_nodes.argtypes = [P(FormDesc), P(BucketDesc), P(ct.c_double)]
        L.dpg_skeleton_create.argtypes = [ct.c_int, ct.c_int, P(ct.c_int32), ct.c_int, ct.c_void_p,
                                          P(ct.c_void_p)]
        L.dpg_skeleton_create_ragged.argtypes = [ct.c_int, P(ct.c_int32), P(ct.c_int32), ct.c_int, ct.c_void_p,
                                                 P(ct.c_void_p)]
        L.dpg_skeleton_info.argtypes = [ct.c_void_p, P(ct.c_int64), P(ct.c_int64), P(ct.c_void_p),
                                        P(ct.c_void_p)]
        L.dpg_skeleton_pattern.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.dpg_skeleton_assemble.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                            ct.c_void_p]
        L.dpg_skeleton_destroy.argtypes = [ct.c_void_p]
        L.dpg_skeleton_solve.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_double,
                                         ct.c_int, P(ct.c_int), P(ct.c_double), ct.c_void_p]
        L.dpg_side_parities.argtypes = [ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.dpg_sort_cells.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.dpg_gather_cells.argtypes = [ct.c_int, ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int,
                                       ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.dpg_coarse_create.argtypes = [P(FormDesc), P(BucketDesc), ct.c_int, P(ct.c_int32), ct.c_void_p,
                                        P(ct.c_void_p)]
        L.dpg_coarse_info.argtypes = [ct.c_void_p, P(ct.c_int64), P(ct.c_int64), P(ct.c_int)]
        L.dpg_coarse_destroy.argtypes = [ct.c_void_p]
        L.dpg_skeleton_solve_twolevel.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                                  ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_int, ct.c_int,
                                                  P(ct.c_int), P(ct.c_double), ct.c_void_p]
        L.dpg_hp_numbering.argtypes = [P(FormDesc), P(BucketDesc), ct.c_int, ct.c_int, P(ct.c_int32), P(ct.c_int32),
                                       ct.c_int, P(ct.c_double), P(ct.c_int64), P(ct.c_int32), P(ct.c_int32),
                                       P(ct.c_int32), P(ct.c_int32), P(ct.c_double)]
        L.dpg_hp_system_create.argtypes = [P(FormDesc), P(BucketDesc), ct.c_int, ct.c_int, P(ct.c_int32),
                                           P(ct.c_int32), ct.c_int, P(ct.c_double), ct.c_void_p, P(ct.c_void_p)]
        L.dpg_hp_system_info.argtypes = [ct.c_void_p, P(ct.c_int64), P(ct.c_int64), P(ct.c_void_p)]
        L.dpg_hp_system_dofs.argtypes = [ct.c_void_p, P(ct.c_int32), P(ct.c_int32), P(ct.c_int32), P(ct.c_int32)]
        L.dpg_hp_system_assemble.argtypes = [ct.c_void_p, P(ct.c_void_p), P(ct.c_void_p), ct.c_void_p, ct.c_void_p,
                                             ct.c_void_p]
        L.dpg_hp_system_local.argtypes = [ct.c_void_p, ct.c_void_p, P(ct.c_void_p), ct.c_void_p]
        L.dpg_hp_system_destroy.argtypes = [ct.c_void_p]
        L.dpg_mark_cells.argtypes = [ct.c_void_p, ct.c_int, ct.c_double, ct.c_void_p, ct.c_void_p]
        L.dpg_riesz_representation.argtypes = [ct.c_void_p, ct.c_int, P(Batch), ct.c_void_p, ct.c_void_p,
                                               ct.c_void_p]
        L.dpg_recover_fields.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                         ct.c_void_p]
        _lib = L
    return _lib


class DPGError(RuntimeError):
    pass


def check(rc):
    if rc < 0:
        raise DPGError(lib().dpg_last_error().decode())
    return rc


def ptr(a):
    """Address of a numpy array or a torch tensor (host or device)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()
