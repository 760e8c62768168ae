"""ctypes mirror of include/gacelle.h and the loader of libgacelle_hip.so.

The structures below follow the C header field by field (tests compile
the header with gcc and compare every size and offset).  The loader never falls back to anything: when the HIP library is missing or
fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int32, c_int64, c_uint8, c_uint32, c_uint64, c_void_p

GC_ABI_VERSION = 9
GC_MAX_PAR = 24
GC_MAX_TABLE = 256
GC_MAX_VOL = 16
GC_MAX_REG = 4
GC_MLP_MAX_LAYERS = 8
GC_MLP_MAX_WIDTH = 80
GC_NOALLOC_MAX_NV = 1 << 25
GC_NOALLOC_SCRATCH_BYTES = 4096 + 8 * (GC_NOALLOC_MAX_NV // 256)

# status codes
GC_OK = 0
GC_ERR_ARG = -1
GC_ERR_HIP = -2
GC_ERR_UNSUPPORTED = -3
GC_ERR_NOMEM = -4
GC_ERR_NWALKER = -5
GC_ERR_INPUT_MISMATCH = -6
GC_ERR_NO_DEVICE = -7

# models
GC_MODEL_R2STAR = 1
GC_MODEL_JOINT_R1R2STAR = 2
GC_MODEL_NEXI = 3
GC_MODEL_MCRMWI = 4

# flags
GC_FLAG_R1R2S_TRUE_FA = 1 << 0
GC_FLAG_NEXI_RICIAN = 1 << 1
GC_FLAG_NEXI_RE_GUARD = 1 << 2
GC_FLAG_MCRMWI_COMPLEX = 1 << 3
GC_FLAG_MCRMWI_EXCHANGE = 1 << 4
GC_FLAG_UNIFORM_TE = 1 << 5
GC_FLAG_MCRMWI_PINI_FA = 1 << 6
GC_FLAG_MCRMWI_EPG = 1 << 7

GC_QUAD_TRAPZ = 0
GC_QUAD_TABLE = 1

GC_OPT_ADAM, GC_OPT_SGDM, GC_OPT_RMSPROP = 0, 1, 2
GC_LOSS_L1, GC_LOSS_L2, GC_LOSS_HUBER, GC_LOSS_MSE = 0, 1, 2, 3
GC_TRANSFORM_SIGMOID, GC_TRANSFORM_LINEAR, GC_TRANSFORM_NONE = 0, 1, 2
GC_CONV_EMA, GC_CONV_LINEAR = 0, 1
GC_ADAM_EFFECTIVE_LR, GC_ADAM_KINGMA_BA = 0, 1
GC_TV_2D, GC_TV_3D = 0, 1
GC_REG_TV, GC_REG_PRIOR, GC_REG_SURFACE_TV = 0, 1, 2
GC_STOP_NONE, GC_STOP_CONVERGENCE, GC_STOP_TOL, GC_STOP_STEP, GC_STOP_GRAD = 0, 1, 2, 3, 4
GC_MCMC_MH, GC_MCMC_GW = 0, 1
GC_ENSEMBLE_SIMULTANEOUS, GC_ENSEMBLE_REDBLACK = 0, 1
GC_RNG_NATIVE, GC_RNG_REPLAY = 0, 1
GC_OPT_GRAPHS, GC_OPT_MCMC_BLOCK, GC_OPT_ASK_BLOCK = 1, 2, 3

STOP_MESSAGES = {
    GC_STOP_NONE: "",
    GC_STOP_CONVERGENCE: "Optimisation is done. Loss convergence below tolerance.",
    GC_STOP_TOL: "Optimisation is done. Loss is less than the tolerance.",
    GC_STOP_STEP: "Optimisation is done. Step norm below tolerance.",
    GC_STOP_GRAD: "Optimisation is done. Gradient norm below tolerance.",
}


class gc_mlp(ctypes.Structure):
    _fields_ = [
        ("n_layer", c_int32),
        ("width", c_int32 * (GC_MLP_MAX_LAYERS + 1)),
        ("alpha", c_float),
        ("reserved", c_int32),
        ("params", c_void_p),
    ]


class gc_model(ctypes.Structure):
    _fields_ = [
        ("model_id", c_int32),
        ("n_meas", c_int32),
        ("n_par", c_int32),
        ("n_extra", c_int32),
        ("flags", c_int32),
        ("i", c_int32 * 8),
        ("f", c_float * 16),
        ("table", c_float * GC_MAX_TABLE),
        ("lb", c_float * GC_MAX_PAR),
        ("ub", c_float * GC_MAX_PAR),
        ("mlp", gc_mlp * 2),
    ]


class gc_askadam_cfg(ctypes.Structure):
    _fields_ = [
        ("iteration", c_int32),
        ("optimiser", c_int32),
        ("loss", c_int32),
        ("transform", c_int32),
        ("adam_form", c_int32),
        ("convergence_model", c_int32),
        ("convergence_window", c_int32),
        ("patience_conv", c_int32),
        ("patience_step", c_int32),
        ("patience_grad", c_int32),
        ("block_iters", c_int32),
        ("tv_mode", c_int32),
        ("n_reg", c_int32),
        ("regmap", c_int32 * GC_MAX_REG),
        ("robust", c_int32),
        ("reg_type", c_int32),
        ("lr0", c_double),
        ("decay", c_double),
        ("beta1", c_double),
        ("beta2", c_double),
        ("eps", c_double),
        ("momentum", c_double),
        ("rms_decay", c_double),
        ("rms_eps", c_double),
        ("tol", c_double),
        ("convergence_value", c_double),
        ("ema_decay", c_double),
        ("step_tol", c_double),
        ("grad_tol", c_double),
        ("lambda_", c_double * GC_MAX_REG),
        ("voxel_size", c_double * 3),
        ("weight_update_interval", c_int32),
        ("outlier_check_window", c_int32),
        ("outlier_weight", c_double),
        ("outlier_voxel_thres", c_double),
        ("outlier_pop_thres", c_double),
        ("outlier_init_thres", c_double),
        ("outlier_init_pop_thres", c_double),
    ]


class gc_volume(ctypes.Structure):
    _fields_ = [
        ("nv", c_int64),
        ("y_dev", c_void_p),
        ("w_dev", c_void_p),
        ("extra_dev", c_void_p),
        ("vox_dev", c_void_p),
        ("dims", c_int32 * 3),
        ("reserved0", c_int32),
        ("z_dev", c_void_p),
        ("zmin_dev", c_void_p),
        ("resloss_dev", c_void_p),
        ("resloss_min_dev", c_void_p),
        ("residual_dev", c_void_p),
        ("residual_min_dev", c_void_p),
        ("loss_trace_host", POINTER(c_double)),
        ("loss", c_double),
        ("loss_fidelity", c_double),
        ("loss_reg", c_double),
        ("min_loss", c_double),
        ("min_loss_fidelity", c_double),
        ("min_loss_reg", c_double),
        ("convergence", c_double),
        ("n_iteration", c_int32),
        ("min_iteration", c_int32),
        ("stop_reason", c_int32),
        ("n_launch", c_int32),
        ("kernel_ms", c_double),
        ("kernel_launches", c_int32),
        ("iters_per_launch", c_int32),
        ("prior_dev", c_void_p),
        ("reg_offset", c_double),
        ("surf_nbr_dev", c_void_p),
        ("surf_len_dev", c_void_p),
        ("surf_max_nb", c_int32),
        ("reserved1", c_int32),
        ("n_mean", c_int64),
        ("sign_trace_dev", c_void_p),
        ("block_threads", c_int32),
    ]


class gc_mcmc_cfg(ctypes.Structure):
    _fields_ = [
        ("algorithm", c_int32),
        ("iteration", c_int32),
        ("thinning", c_int32),
        ("n_burnin", c_int32),
        ("repetition", c_int32),
        ("n_walker", c_int32),
        ("ensemble_update", c_int32),
        ("rng_mode", c_int32),
        ("noise_slot", c_int32),
        ("n_var", c_int32),
        ("seed", c_uint64),
        ("step_size", c_double),
        ("start_range", c_double),
        ("x_step", c_float * GC_MAX_PAR),
        ("global_mask", c_uint32),
        ("global_warmup", c_int32),
    ]


class gc_mcmc_io(ctypes.Structure):
    _fields_ = [
        ("nv", c_int64),
        ("y_dev", c_void_p),
        ("w_dev", c_void_p),
        ("extra_dev", c_void_p),
        ("x0_dev", c_void_p),
        ("normals_dev", c_void_p),
        ("uniforms_dev", c_void_p),
        ("partners_dev", c_void_p),
        ("g_uniforms_dev", c_void_p),
        ("g_partners_dev", c_void_p),
        ("posterior_dev", c_void_p),
        ("mean_dev", c_void_p),
        ("std_dev", c_void_p),
        ("accept_dev", c_void_p),
        ("xfinal_dev", c_void_p),
        ("trace_dev", c_void_p),
        ("g_posterior_dev", c_void_p),
        ("g_xfinal_dev", c_void_p),
        ("g_trace_dev", c_void_p),
        ("n_sample", c_int32),
        ("n_stored", c_int32),
        ("kernel_ms", c_double),
        ("block_threads", c_int32),
    ]


# exported symbols and their signatures (argtypes, restype)
SIGNATURES = {
    "gc_version": ([], c_int32),
    "gc_ctx_create": ([c_int32, POINTER(c_void_p)], c_int32),
    "gc_ctx_destroy": ([c_void_p], c_int32),
    "gc_ctx_set_stream": ([c_void_p, c_void_p], c_int32),
    "gc_ctx_synchronize": ([c_void_p], c_int32),
    "gc_ctx_set_option": ([c_void_p, c_int32, c_int64], c_int32),
    "gc_last_error": ([c_void_p], ctypes.c_char_p),
    "gc_nd_to_gd": ([c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_int64, c_void_p], c_int32),
    "gc_gd_to_nd": ([c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_int64, c_void_p], c_int32),
    "gc_rescale_parameters": ([c_void_p, POINTER(gc_model), c_int32, c_void_p, c_void_p, c_int64], c_int32),
    "gc_unscale_parameters": ([c_void_p, POINTER(gc_model), c_int32, c_void_p, c_void_p, c_int64], c_int32),
    "gc_fwd": ([c_void_p, POINTER(gc_model), c_void_p, c_void_p, c_void_p, c_int64], c_int32),
    "gc_fwd_adj": ([c_void_p, POINTER(gc_model), c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                    c_int64, c_float, c_void_p, c_void_p, c_void_p], c_int32),
    "gc_fwd_adj_noalloc": ([c_void_p, POINTER(gc_model), c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_int64, c_float, c_void_p, c_void_p, c_void_p], c_int32),
    "gc_tv_fwd_adj": ([c_void_p, c_void_p, c_void_p, POINTER(c_int32), c_int32, POINTER(c_double), c_double,
                       c_int64, c_void_p, c_void_p], c_int32),
    "gc_posterior_stats": ([c_void_p, c_void_p, c_int32, c_int32, c_int64, POINTER(c_float), POINTER(c_float), c_int32,
                            c_void_p, c_void_p, c_void_p], c_int32),
    "gc_dict_search": ([c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_int64, POINTER(c_float), c_void_p,
                        c_void_p], c_int32),
    "gc_askadam_run": ([c_void_p, POINTER(gc_model), POINTER(gc_askadam_cfg), POINTER(gc_volume), c_int32],
                       c_int32),
    "gc_mcmc_run": ([c_void_p, POINTER(gc_model), POINTER(gc_mcmc_cfg), POINTER(gc_mcmc_io)], c_int32),
}

LIB_NAME = "libgacelle_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

_lib = None


class GacelleError(RuntimeError):
    """Error raised by a library call (the MATLAB error identifiers' role)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


def load_library(path: str | None = None):
    """Load libgacelle_hip.so (built in-tree by __graft_entry__.build)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise GacelleError(GC_ERR_NO_DEVICE, f"HIP library not built: {p} is missing (run __graft_entry__.build())")
    lib = ctypes.CDLL(p)
    for name, (argtypes, restype) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    if lib.gc_version() != GC_ABI_VERSION:
        raise GacelleError(GC_ERR_ARG, f"ABI version mismatch: library {lib.gc_version()} vs {GC_ABI_VERSION}")
    if path is None:
        _lib = lib
    return lib
