"""ctypes binding of ``libmrirecon_hip.so`` (declared in include/mrirecon_hip.h).

The library is loaded after ``import torch`` so that the process keeps a
single HIP runtime (torch's bundled ``libamdhip64.so.7``).  There is no CPU
fallback: if the library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_double, c_float, c_int, c_int64, c_void_p
from pathlib import Path

import torch  # noqa: F401  (loads the HIP runtime first)

# MRH_LIB selects another build of the same ABI (e.g. a debug build); default is the in-tree one.
LIB_PATH = Path(os.environ.get("MRH_LIB") or Path(__file__).resolve().parent / "libmrirecon_hip.so")
ABI_VERSION = 13

DCF_NONE, DCF_BALANCED, DCF_STANDARD = 0, 1, 2
TV_AXIS_T, TV_AXIS_X, TV_AXIS_Y = 1, 2, 4
CG_DIR_FR, CG_DIR_PRPLUS, CG_DIR_DY = 0, 1, 2
CG_LS_ARMIJO, CG_LS_WOLFE = 0, 1
CG_ACTIVE, CG_LS_FAILED, CG_NOT_DESCENT, CG_REL_TOL, CG_TARGET = 0, 1, 2, 3, 4


class NufftDesc(Structure):
    _fields_ = [
        ("ndim", c_int), ("S0", c_int), ("S1", c_int), ("G0", c_int), ("G1", c_int), ("W", c_int),
        ("beta", c_double), ("C", c_int), ("B", c_int), ("like", c_int),
        ("k_offsets", POINTER(c_int64)), ("traj", POINTER(c_float)), ("dcf", POINTER(c_float)),
        ("dcf_mode", c_int), ("dcf_in_fwd", c_int), ("dcf_in_adj", c_int),
        ("maps", c_void_p), ("strip_cols", c_int),
    ]


class Nufft3Desc(Structure):
    _fields_ = [
        ("S0", c_int), ("S1", c_int), ("S2", c_int), ("G0", c_int), ("G1", c_int), ("G2", c_int), ("W", c_int),
        ("beta", c_double), ("C", c_int), ("B", c_int),
        ("k_offsets", POINTER(c_int64)), ("traj", POINTER(c_float)), ("dcf", POINTER(c_float)),
        ("dcf_mode", c_int), ("dcf_in_fwd", c_int), ("dcf_in_adj", c_int), ("maps", c_void_p),
    ]


class NufftInfo(Structure):
    _fields_ = [
        ("G0", c_int), ("G1", c_int), ("W", c_int), ("strip_cols", c_int), ("n_strips", c_int),
        ("kb_terms", c_int), ("beta", c_double), ("n_samples", c_int64), ("n_entries", c_int64),
        ("max_strip_entries", c_int64), ("workspace_bytes", c_int64), ("coil_chunk", c_int),
    ]


class TVTerm(Structure):
    _fields_ = [("weight", c_float), ("eps", c_float), ("axes", c_int), ("isotropic", c_int),
                ("axis_w", c_float * 3)]


class TVDesc(Structure):
    _fields_ = [("P", c_int), ("F", c_int), ("S0", c_int), ("S1", c_int), ("n_terms", c_int),
                ("term", TVTerm * 3)]


class CGDesc(Structure):
    _fields_ = [("tv", TVDesc), ("k_prob", POINTER(c_int64)), ("direction", c_int), ("c1", c_double),
                ("ls_beta", c_double), ("ls_max_iter", c_int), ("rel_tol", c_double), ("max_hist", c_int),
                ("restart_c", c_double), ("ls_kind", c_int), ("c2", c_double),
                ("diag", c_void_p), ("gl_weight", c_double), ("gl_F", c_int), ("gl_L", c_void_p),
                ("gl_prof", c_void_p), ("fp32_reductions", c_int)]


class CGHist(Structure):
    _fields_ = [(n, c_double) for n in ("f0", "g0d", "t", "ft", "rho", "beta", "freg", "gdot_step",
                                        "xnorm2", "dnorm2")]


_SIGS = {
    "mrh_last_error": (ctypes.c_char_p, []),
    "mrh_abi_version": (c_int, []),
    "mrh_nufft_plan_create": (c_int, [POINTER(NufftDesc), POINTER(c_void_p)]),
    "mrh_nufft_plan_destroy": (c_int, [c_void_p]),
    "mrh_nufft_plan_info": (c_int, [c_void_p, POINTER(NufftInfo)]),
    "mrh_nufft_set_coil_chunk": (c_int, [c_void_p, c_int]),
    "mrh_nufft_forward": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "mrh_nufft_adjoint": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "mrh_nufft_forward_masked": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_void_p]),
    "mrh_nufft_adjoint_masked": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p]),
    "mrh_nufft_forward_stage": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "mrh_nufft_adjoint_stage": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p]),
    "mrh_nufft3_plan_create": (c_int, [POINTER(Nufft3Desc), POINTER(c_void_p)]),
    "mrh_nufft3_plan_destroy": (c_int, [c_void_p]),
    "mrh_nufft3_forward": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "mrh_nufft3_adjoint": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "mrh_reduce_workspace_len": (c_int64, []),
    "mrh_residual_energy": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_int64,
                                    c_void_p, c_void_p, c_void_p]),
    "mrh_dot_real": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "mrh_tv_energy_grad": (c_int, [POINTER(TVDesc), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mrh_tv_energy_grad_noalloc": (c_int, [POINTER(TVDesc), c_void_p, c_void_p, c_void_p, c_void_p]),
    "mrh_cg_create": (c_int, [POINTER(CGDesc), POINTER(c_void_p)]),
    "mrh_cg_destroy": (c_int, [c_void_p]),
    "mrh_cg_reset": (c_int, [c_void_p, c_void_p]),
    "mrh_cg_residual": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mrh_cg_linesearch": (c_int, [c_void_p] + [c_void_p] * 6 + [c_void_p]),
    "mrh_cg_accept": (c_int, [c_void_p] + [c_void_p] * 6 + [c_void_p]),
    "mrh_cg_gradient": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "mrh_cg_direction": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "mrh_cg_set_targets": (c_int, [c_void_p, POINTER(c_double), c_void_p]),
    "mrh_cg_active_flags": (c_int, [c_void_p, POINTER(c_void_p)]),
    "mrh_cg_build_diag": (c_int, [c_void_p, c_void_p, c_void_p]),
    "mrh_cg_read_status": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), c_void_p]),
    "mrh_cg_read_history": (c_int, [c_void_p, POINTER(CGHist), c_int, c_void_p]),
    "mrh_cg_read_latest": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(CGHist), c_void_p]),
    "mrh_unet_conv_fwd": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p] + [c_int] * 9 + [c_void_p]),
    "mrh_unet_conv_bwd_weight": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p] + [c_int] * 8 + [c_void_p]),
    "mrh_unet_relu_bwd": (c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    "mrh_unet_pool2": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p]),
    "mrh_unet_pool2_bwd": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p]),
    "mrh_unet_up2_concat": (c_int, [c_void_p, c_void_p, c_void_p] + [c_int] * 5 + [c_void_p]),
    "mrh_unet_conv3x3_upcat": (c_int, [c_void_p] * 5 + [c_int] * 9 + [c_void_p]),
    "mrh_unet_upcat_supported": (c_int, [c_int] * 4),
    "mrh_unet_up2_concat_bwd": (c_int, [c_void_p, c_void_p, c_void_p] + [c_int] * 5 + [c_void_p]),
    "mrh_unet_pack_input": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_float, c_void_p]),
    "mrh_unet_unpack_residual": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_float, c_void_p]),
    "mrh_unet_level0_fused_supported": (c_int, [c_int, c_int]),
    "mrh_unet_enc0_fused": (c_int, [c_void_p, c_float] + [c_void_p] * 6 + [c_int] * 3 + [c_void_p]),
    "mrh_unet_dec0_tail": (c_int, [c_void_p] * 6 + [c_float, c_void_p] + [c_int] * 3 + [c_void_p]),
    "mrh_unet_mse_grad": (c_int, [c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_void_p, c_int64, c_int,
                                  c_float, c_float, c_void_p]),
    "mrh_unet_fill": (c_int, [c_void_p, c_int64, c_float, c_void_p]),
    "mrh_unet_adam": (c_int, [c_void_p] * 10 + [c_int] * 4 + [c_float] * 4 + [c_int, c_void_p]),
    "mrh_tbasis_forward": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, c_void_p, c_void_p]),
    "mrh_tbasis_adjoint": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, c_void_p, c_int, c_void_p]),
    "mrh_tbasis_diag": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_int64, c_void_p, c_void_p]),
}

_ERR = {1: ValueError, 2: RuntimeError, 3: MemoryError, 4: NotImplementedError}
_lib = None


def exported_symbols():
    """Names declared by include/mrirecon_hip.h that the library must export."""
    return sorted(_SIGS)


def load() -> ctypes.CDLL:
    """Load (once) and type the C-ABI library; raises if it is missing or stale."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m mrirecon_amd.build` "
            "(there is no CPU fallback for the HIP path)")
    lib = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    v = lib.mrh_abi_version()
    if v != ABI_VERSION:
        raise ImportError(f"{LIB_PATH} has ABI {v}, expected {ABI_VERSION}: rebuild it")
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().mrh_last_error().decode(errors="replace")
        raise _ERR.get(rc, RuntimeError)(msg)


def stream_of(t: torch.Tensor) -> int:
    """Handle of the stream that is current for ``t``'s device (never cached)."""
    return torch.cuda.current_stream(t.device).cuda_stream


def ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())
