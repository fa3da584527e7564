"""ctypes binding of libbfgpu.so (declarations in include/bfgpu.h).

The library is built in-tree (``butterflyfactorizations.jl_amd/lib``); loading
fails loudly if it is missing -- there is no CPU fallback for any entry point.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libbfgpu.so")

BFGPU_C64, BFGPU_C128 = 1, 2
OP_N, OP_T, OP_C = 0, 1, 2
TREE_BALANCED, TREE_MAXPOINTS = 0, 1
RANK_BUTTERFLY, RANK_GEOMETRIC, RANK_LEARNED = 0, 1, 2
XY_HOST, XY_DEVICE = 0, 1

i32, i64, u64, f32, f64, vp = C.c_int32, C.c_int64, C.c_uint64, C.c_float, C.c_double, C.c_void_p
P = C.POINTER


class TreeView(C.Structure):
    _fields_ = [
        ("n_nodes", i32), ("n_values", i32),
        ("parent", P(i32)), ("first_child", P(i32)), ("n_children", P(i32)),
        ("val_begin", P(i32)), ("val_end", P(i32)),
        ("center", P(f64)), ("halfsize", P(f64)), ("perm", P(i32)),
        ("level", P(i32)), ("points", P(f64)),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("k", f64), ("tol", f64), ("adaptive", i32), ("rank_mode", i32),
        ("c_tau", f64), ("rmin", i32), ("c_geo", f64), ("c_eps", f64),
        ("precision", i32), ("oversample", i32), ("seed", u64), ("mlp", vp), ("workspace_bytes", i64),
        ("kernel", vp), ("judge_rows", i32), ("judge_slack", f64),
    ]


class BFInfo(C.Structure):
    _fields_ = [
        ("n_bf", i32), ("precision", i32), ("rows", i64), ("cols", i64),
        ("max_stages", i32), ("n_blocks", i64), ("n_data", i64),
        ("factor_bytes", i64), ("slot_len", i64), ("n_comp", i64),
        ("n_retry", i64), ("t_factor_ms", f64),
        ("t_plan_ms", f64), ("t_sample_ms", f64), ("t_qr_ms", f64), ("t_interp_ms", f64), ("t_pack_ms", f64),
        ("device_bytes", i64), ("n_judged", i64), ("n_fallback", i64),
    ]


class BfgpuError(RuntimeError):
    """A failed libbfgpu call; carries the status code and message."""

    def __init__(self, code, msg):
        super().__init__(f"bfgpu error {code}: {msg}")
        self.code = code


class ArgumentError(BfgpuError, ValueError):
    """Invalid argument (the reference raises ArgumentError/DimensionMismatch)."""


_SIGS = {
    "bfgpu_last_error": (C.c_char_p, []),
    "bfgpu_version": (C.c_char_p, []),
    "bfgpu_ctx_create": (i32, [C.c_int, P(vp)]),
    "bfgpu_ctx_destroy": (i32, [vp]),
    "bfgpu_ctx_set_stream": (i32, [vp, vp]),
    "bfgpu_ctx_reset_stream": (i32, [vp]),
    "bfgpu_bf_save": (i32, [vp, C.c_char_p]),
    "bfgpu_bf_recompress": (i32, [vp, f64, P(vp)]),
    "bfgpu_bf_materialize": (i32, [vp, i32, P(vp)]),
    "bfgpu_bf_load": (i32, [vp, vp, vp, C.c_char_p, P(vp)]),
    "bfgpu_ctx_get_stream": (i32, [vp, P(vp)]),
    "bfgpu_ctx_sync": (i32, [vp]),
    "bfgpu_tree_build": (i32, [P(f64), i64, i32, i32, i32, P(vp)]),
    "bfgpu_tree_create": (i32, [P(f64), i32, i32, P(i32), P(i32), P(i32), P(i32), P(i32),
                                P(f64), P(f64), P(i32), P(vp)]),
    "bfgpu_tree_view_get": (i32, [vp, P(TreeView)]),
    "bfgpu_tree_destroy": (i32, [vp]),
    "bfgpu_nearandfar": (i32, [vp, vp, i32, f64, i32, i32, i32, i32, P(i32), P(i64), P(i32), P(i64)]),
    "bfgpu_factorize": (i32, [vp, vp, vp, i32, P(i32), P(Opts), P(vp)]),
    "bfgpu_kernel_rwg_create": (i32, [vp, P(f64), P(f64), i64, i32, P(f64), i64, i32, P(f64), P(vp)]),
    "bfgpu_kernel_fill": (i32, [vp, f64, i32, i32, P(i32), i32, P(i32), vp]),
    "bfgpu_kernel_destroy": (i32, [vp]),
    "bfgpu_kernel_rwg_set_near": (i32, [vp, P(f64), P(f64), i32, i32, i32, i32, f64]),
    "bfgpu_kernel_fill_near": (i32, [vp, f64, i32, i32, P(i32), i32, P(i32), vp]),
    "bfgpu_near_assemble_kernel": (i32, [vp, vp, vp, vp, f64, i32, i64, P(i32), P(vp)]),
    "bfgpu_near_assemble": (i32, [vp, vp, vp, f64, i32, P(f64), i64, P(i32), P(vp)]),
    "bfgpu_near_matvec": (i32, [vp, i32, i64, vp, i64, vp, i64, P(f64), P(f64), i32]),
    "bfgpu_near_info": (i32, [vp, P(i64)]),
    "bfgpu_near_block": (i32, [vp, i64, P(i32), vp]),
    "bfgpu_near_set_transposed_copy": (i32, [vp, i32]),
    "bfgpu_near_set_mfma": (i32, [vp, i32]),
    "bfgpu_near_destroy": (i32, [vp]),
    "bfgpu_bf_from_host": (i32, [vp, vp, vp, i32, i64, P(i32), P(i64), P(f64), i32, P(vp)]),
    "bfgpu_bf_info_get": (i32, [vp, P(BFInfo)]),
    "bfgpu_bf_counts": (i32, [vp, i32, P(i32), P(i64), P(i64), P(i64), P(i64), P(i64), P(i32)]),
    "bfgpu_bf_export": (i32, [vp, i32, P(i32), P(i64), P(f64)]),
    "bfgpu_bf_export_logs": (i32, [vp, i32, P(i32), P(i32), P(i32)]),
    "bfgpu_bf_export_features": (i32, [vp, i32, P(f32)]),
    "bfgpu_bf_export_judge": (i32, [vp, i32, P(i32)]),
    "bfgpu_bf_compressions": (i32, [vp, P(i64), P(i32), P(f32)]),
    "bfgpu_bf_stage_stats": (i32, [vp, P(i32), P(i64)]),
    "bfgpu_bf_launch_stats": (i32, [vp, i32, P(i32), P(i64)]),
    "bfgpu_bf_stream_stats": (i32, [vp, P(i32), P(i64)]),
    "bfgpu_bf_convert": (i32, [vp, i32, P(vp)]),
    "bfgpu_matvec": (i32, [vp, i32, i64, vp, i64, vp, i64, P(f64), P(f64), i32]),
    "bfgpu_matvec_noalloc": (i32, [vp, i32, i64, vp, i64, vp, i64, P(f64), P(f64)]),
    "bfgpu_matvec_timed": (i32, [vp, i32, i64, vp, i64, vp, i64, i32, P(f64), P(i64), P(i32), P(f64)]),
    "bfgpu_matvec_launch_times": (i32, [vp, i32, i64, vp, i64, vp, i64, i32, i32, P(i32), P(f64), P(i32)]),
    "bfgpu_bf_get_tuning": (i32, [vp, P(i32), P(i32)]),
    "bfgpu_bf_set_tuning": (i32, [vp, i32, P(i32)]),
    "bfgpu_bf_set_fused": (i32, [vp, i32]),
    "bfgpu_bf_set_option": (i32, [vp, i32, i32]),
    "bfgpu_bf_get_option": (i32, [vp, i32, P(i32)]),
    "bfgpu_bf_check": (i32, [vp]),
    "bfgpu_bf_autotune": (i32, [vp, i64, i32, P(f64), P(i32)]),
    "bfgpu_bf_destroy": (i32, [vp]),
    "bfgpu_mlp_create": (i32, [vp, i32, i32, i32, u64, P(vp)]),
    "bfgpu_mlp_destroy": (i32, [vp]),
    "bfgpu_mlp_n_params": (i32, [vp, P(i64)]),
    "bfgpu_mlp_get_params": (i32, [vp, P(f32)]),
    "bfgpu_mlp_set_params": (i32, [vp, P(f32)]),
    "bfgpu_mlp_set_norm": (i32, [vp, P(f32), P(f32), f32, f32]),
    "bfgpu_mlp_forward": (i32, [vp, i64, P(f32), P(f32)]),
    "bfgpu_mlp_loss_grad": (i32, [vp, i64, P(f32), P(f32), P(f32), P(f32)]),
    "bfgpu_mlp_train": (i32, [vp, i64, P(f32), P(f32), i32, i32, f32, u64, P(f32)]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def lib():
    """Load libbfgpu.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        # One HIP runtime per process: torch ships its own libamdhip64 under
        # a different file name, so load torch first (when present) and let
        # libbfgpu bind to the runtime it brought in.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"libbfgpu.so not found at {LIB_PATH}; build it with "
                "butterflyfactorizations.jl_amd/build.sh (or __graft_entry__.build())")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def check(status):
    if status != 0:
        msg = lib().bfgpu_last_error().decode(errors="replace")
        if status == 1:
            raise ArgumentError(status, msg)
        raise BfgpuError(status, msg)


def ptr(a, ctype):
    """Pointer to a contiguous numpy array's data."""
    return a.ctypes.data_as(P(ctype))
