"""ctypes binding of libvipe_hip.so (the C ABI declared in include/vipe_hip.h).

The library is built in-tree by ``python -m vipe_amd.build``.  There is no
fallback: if the shared object is missing or fails to load, every op raises.
``torch`` is imported first so the HIP runtime that torch ships is the one the
library binds to (both carry the soname libamdhip64.so.7).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import torch  # noqa: F401  (loads the HIP runtime before the library)

LIB_PATH = Path(os.environ.get("VIPE_HIP_LIB", Path(__file__).resolve().parent / "_lib" / "libvipe_hip.so"))

VH_OK, VH_ERR_ARG, VH_ERR_HIP, VH_ERR_UNSUPPORTED, VH_ERR_UNDEFINED = 0, 1, 2, 3, 4
VH_F16, VH_F32, VH_F64, VH_BF16, VH_FP8E4M3 = 1, 2, 3, 4, 5
VH_U8, VH_I8, VH_I16, VH_I32, VH_I64 = 6, 7, 8, 9, 10
VH_LAYOUT_PLANAR, VH_LAYOUT_TILED = 0, 1
VH_DEVICE, VH_HOST = 0, 1

vp = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
f32 = C.c_float
sz = C.c_size_t

# name -> (restype, argtypes)
_SIGS = {
    "vh_last_error": (C.c_char_p, []),
    "vh_abi_version": (i32, []),
    "vh_fill_i32": (i32, [vp, i64, i32, vp]),
    # lietorch
    "vh_lie_expm": (i32, [i32, i32, i64, vp, vp, i32, vp]),
    "vh_lie_expm_backward": (i32, [i32, i32, i64, vp, vp, vp, i32, vp]),
    "vh_lie_logm": (i32, [i32, i32, i64, vp, vp, i32, vp]),
    "vh_lie_logm_backward": (i32, [i32, i32, i64, vp, vp, vp, i32, vp]),
    "vh_lie_inv": (i32, [i32, i32, i64, vp, vp, i32, vp]),
    "vh_lie_inv_backward": (i32, [i32, i32, i64, vp, vp, vp, i32, vp]),
    "vh_lie_mul": (i32, [i32, i32, i64, vp, vp, vp, i32, vp]),
    "vh_lie_mul_backward": (i32, [i32, i32, i64, vp, vp, vp, vp, vp, i32, vp]),
    "vh_lie_adj": (i32, [i32, i32, i64, vp, vp, vp, i32, vp]),
    "vh_lie_adj_backward": (i32, [i32, i32, i64, vp, vp, vp, vp, vp, i32, vp]),
    "vh_lie_adjT": (i32, [i32, i32, i64, vp, vp, vp, i32, vp]),
    "vh_lie_adjT_backward": (i32, [i32, i32, i64, vp, vp, vp, vp, vp, i32, vp]),
    "vh_lie_act": (i32, [i32, i32, i64, vp, vp, vp, i32, vp]),
    "vh_lie_act_backward": (i32, [i32, i32, i64, vp, vp, vp, vp, vp, i32, vp]),
    "vh_lie_act4": (i32, [i32, i32, i64, vp, vp, vp, i32, vp]),
    "vh_lie_act4_backward": (i32, [i32, i32, i64, vp, vp, vp, vp, vp, i32, vp]),
    "vh_lie_as_matrix": (i32, [i32, i32, i64, vp, vp, i32, vp]),
    "vh_lie_projector": (i32, [i32, i32, i64, vp, vp, i32, vp]),
    "vh_lie_Jinv": (i32, [i32, i32, i64, vp, vp, vp, i32, vp]),
    # correlation
    "vh_corr_index_forward": (i32, [i32, i64, i32, i32, i32, i32, vp, vp, i32, vp, vp]),
    "vh_corr_index_backward": (i32, [i32, i64, i32, i32, i32, i32, vp, vp, i32, vp, vp]),
    "vh_corr_lookup_pyramid": (i32, [i32, i64, i32, i32, i32, vp, vp, vp, i32, vp, i32, vp, vp]),
    "vh_corr_build_workspace": (sz, [i32, i64, i32, i32, i32]),
    "vh_corr_build": (i32, [i32, i32, i32, i64, i32, i32, i32, vp, vp, i32, vp, i32, vp, vp]),
    "vh_altcorr_forward": (i32, [i32, i64, i32, i32, i32, i32, i32, i32, vp, vp, vp, i32, vp, vp]),
    "vh_altcorr_backward": (i32, [i32, i64, i32, i32, i32, i32, i32, i32, vp, vp, vp, vp, i32, vp, vp, vp, vp]),
    "vh_feat_to_nhwc": (i32, [i32, vp, i64, i32, i32, f32, i32, vp, vp]),
    "vh_pool2x2_nhwc": (i32, [i32, vp, i64, i32, i32, i32, vp, vp]),
    "vh_corr_feature_grad": (i32, [i32, i64, i32, i32, i32, i32, vp, vp, f32, vp, vp, vp]),
    "vh_altcorr_backward_workspace_size": (C.c_size_t, [i64, i32, i32, i32]),
    "vh_altcorr_index_forward": (i32, [i32, i64, i32, i32, i32, i32, i32, vp, i32, vp, vp, vp, vp, vp, vp, i32,
                                       vp, vp]),
    # bundle adjustment
    "vh_ba_plan_create": (vp, [vp, vp, i32, i32, i32, i32, i32, i32]),
    "vh_ba_plan_create_tiled": (vp, [vp, vp, i32, i32, i32, i32, i32, i32, i32]),
    "vh_ba_plan_pixel_tile": (i32, [vp]),
    "vh_ba_plan_destroy": (None, [vp]),
    "vh_ba_plan_num_depth_frames": (i32, [vp]),
    "vh_ba_plan_depth_frames": (i32, [vp, vp]),
    "vh_ba_workspace_size": (sz, [vp]),
    "vh_ba_adjoint": (i32, [vp, vp, vp, vp, vp, vp, vp, f32, i32, vp, vp, vp, vp, vp, vp]),
    "vh_ba_run": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, f32, f32, f32, f32, f32, f32, i32, vp, vp, vp, vp,
                        vp]),
    "vh_ba_run_static": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, f32, f32, f32, f32, f32, f32, i32, vp, vp,
                               vp, vp]),
    "vh_ba_run_robust": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, f32, f32, f32, f32, f32, f32, i32, i32, f32,
                               f32, f32, vp, vp, vp, vp, vp]),
    "vh_robust_weights": (i32, [i32, f32, f32, f32, vp, vp, i64, vp]),
    "vh_ba_plan_create_rig": (vp, [vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, i32, i32]),
    "vh_ba_run_rig": (i32, [vp, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, i32, f32, f32, f32, i32, i32, f32, f32,
                            f32, vp, vp, vp, vp, vp]),
    # geometry siblings
    "vh_geom_iproj": (i32, [vp, vp, vp, i32, i32, i32, vp, vp]),
    "vh_geom_projmap": (i32, [vp, vp, vp, vp, vp, i32, i32, i32, i32, vp, vp, vp]),
    "vh_geom_frame_distance": (i32, [vp, vp, vp, i32, vp, vp, vp, vp, vp, i32, i32, i32, i32, f32, vp, vp]),
    "vh_geom_depth_filter": (i32, [vp, vp, vp, vp, vp, i32, i32, i32, i32, vp, vp]),
    # corr_ext spatial correlation sampler
    "vh_spatial_corr_out_size": (i32, [i32] * 10 + [C.POINTER(i32), C.POINTER(i32)]),
    "vh_spatial_corr_workspace_size": (sz, [i32] * 17),
    "vh_spatial_corr_forward": (i32, [i32, vp, vp] + [i32] * 16 + [vp, vp, sz, vp]),
    "vh_spatial_corr_backward_workspace_size": (sz, [i32] * 17),
    "vh_spatial_corr_backward": (i32, [i32, vp, vp, vp] + [i32] * 16 + [vp, vp, vp, sz, vp]),
    # factor-graph update glue
    "vh_graph_motion": (i32, [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, vp, vp, vp]),
    "vh_graph_targets": (i32, [i32, vp, vp, vp, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp, vp]),
    # DroidNet update operator
    "vh_uo_conv_args_size": (sz, []),
    "vh_uo_conv": (i32, [vp, vp]),
    "vh_uo_conv_small": (i32, [vp, i32, i32, i32, vp, vp, i32, i32, i32, i32, i32, i32, i32, vp, i32, vp]),
    "vh_uo_to_nhwc": (i32, [i32, vp, i32, i32, i32, i32, vp, vp]),
    "vh_uo_to_nchw": (i32, [i32, vp, i32, i32, i32, i32, i32, vp, vp]),
    "vh_uo_mean": (i32, [vp, i32, i32, i32, vp, vp]),
    "vh_uo_fc": (i32, [vp, i32, i32, vp, vp, i32, vp, vp]),
    # scatter_ext
    "vh_scatter_workspace_size": (sz, [i64] * 7),
    "vh_scatter": (i32, [i32, i32, vp, vp] + [i64] * 7 + [vp, vp, i32, vp, sz, vp]),
}



class UoConvArgs(C.Structure):
    """vh_uo_conv_args (include/vipe_hip.h)."""
    _fields_ = [("seg", vp * 4), ("seg_c", i32 * 4), ("seg_stride", i32 * 4), ("seg_off", i32 * 4), ("nseg", i32),
                ("weight", vp), ("bias", vp), ("nimg", i32), ("H", i32), ("W", i32), ("Cout", i32),
                ("Cout_pad", i32), ("KH", i32), ("KW", i32), ("epilogue", i32), ("out", vp), ("out_f32", i32),
                ("out_stride", i32), ("out_off", i32), ("glo", vp), ("glo_stride", i32), ("aux0", vp), ("aux1", vp),
                ("aux_stride", i32), ("out2", vp)]


VH_UO_NONE, VH_UO_RELU, VH_UO_SIGMOID, VH_UO_ETA, VH_UO_SIGMUL, VH_UO_GATES, VH_UO_GRU = 0, 1, 2, 3, 4, 5, 6
VH_UO_NCHW = 16

_lib = None
_load_error = None


def exported_symbols():
    return list(_SIGS)


def load():
    """Load (once) and return the ctypes library; raises if it is unavailable."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"libvipe_hip.so not found at {LIB_PATH}; build it with `python -m vipe_amd.build` "
            "(there is no non-native fallback)."
        )
    lib = C.CDLL(str(LIB_PATH))
    missing = []
    for name, (res, args) in _SIGS.items():
        try:
            fn = getattr(lib, name)
        except AttributeError:
            missing.append(name)
            continue
        fn.restype = res
        fn.argtypes = args
    if missing:
        raise RuntimeError(f"libvipe_hip.so is missing symbols: {missing}")
    _lib = lib
    return lib


def last_error() -> str:
    return load().vh_last_error().decode(errors="replace")


class VHUndefined(Exception):
    """The reference op would return an undefined tensor (None in Python)."""


def check(rc: int, what: str):
    if rc == VH_OK:
        return
    if rc == VH_ERR_UNDEFINED:
        raise VHUndefined(last_error())
    raise RuntimeError(f"{what}: {last_error()}")


def ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def stream_of(t):
    if t.device.type == "cpu":
        return None
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def place_of(t):
    return VH_HOST if t.device.type == "cpu" else VH_DEVICE


DTYPE_CODE = {torch.float16: VH_F16, torch.float32: VH_F32, torch.float64: VH_F64, torch.bfloat16: VH_BF16,
              torch.uint8: VH_U8, torch.int8: VH_I8, torch.int16: VH_I16, torch.int32: VH_I32, torch.int64: VH_I64}


def dtype_code(t):
    return DTYPE_CODE.get(t.dtype, -1)
