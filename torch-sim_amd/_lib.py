"""ctypes binding of the tsx C ABI (include/tsx.h).

The library is loaded after ``torch`` so that it resolves libamdhip64.so.7 to
the HIP runtime PyTorch already loaded: device pointers and hipStream_t
handles from torch are then valid in the library, and kernels it launches on
a capturing torch stream become nodes of the torch CUDA/HIP graph.

There is no fallback: if libtsx.so is missing or the device is not a ROCm GPU
the calls raise.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import torch

_LIB_PATH = Path(__file__).resolve().parent / "libtsx.so"

P, I, L, D, U = (
    ctypes.c_void_p,
    ctypes.c_int32,
    ctypes.c_int64,
    ctypes.c_double,
    ctypes.c_uint64,
)

# name -> (restype, argtypes); mirrors include/tsx.h
_SIGS = {
    "tsx_version": (I, []),
    "tsx_last_error": (ctypes.c_char_p, []),
    "tsx_nl_ws_bytes": (L, [L, I, I]),
    "tsx_nl_build": (
        I,
        [P, I, P, P, I, P, P, L, I, D, I, I, I, I, I, I, L, I, I, P, P, P, P, P, P, P, P, P, I, I,
         P, I, P],
    ),
    "tsx_nl_build_noalloc": (
        I,
        [P, I, P, P, I, P, P, L, I, D, I, I, I, I, I, I, L, I, P, P, P, P, P, P, P, P, I, I, P, I,
         P],
    ),
    "tsx_nl_check_ws_bytes": (L, [I]),
    "tsx_nl_check": (I, [P, I, P, P, P, P, L, I, D, D, P, P, P]),
    "tsx_scan_ws_bytes": (L, [L]),
    "tsx_nl_count_pairs": (I, [P, L, P, P, P]),
    "tsx_nl_emit_pairs": (I, [P, P, P, P, P, P, L, I, I, L, P, L, P, P, P]),
    "tsx_wrap_ws_bytes": (L, [I]),
    "tsx_wrap_positions": (I, [P, P, P, I, P, L, I, P, P, P, P, P]),
    "tsx_pair_forces": (
        I,
        [I, P, I, P, P, I, P, P, P, P, P, P, L, I, I, L, D, I, P, P, P, P],
    ),
    "tsx_system_reduce": (I, [P, P, P, P, I, I, P, P, P, P, P]),
    "tsx_reduce_ws_bytes": (L, [I, I]),
    "tsx_md_kick": (I, [P, P, D, L, P]),
    "tsx_md_drift": (I, [P, P, P, D, L, P]),
    "tsx_md_vv_half": (I, [P, P, P, P, D, L, P]),
    "tsx_md_baoa": (I, [P, P, P, P, P, P, P, D, P, U, U, L, P]),
    "tsx_md_fused_step": (
        I,
        [
            I, P, I, P, P, P, P, P, P, P, P, P, L, I, I, L, D,  # pair part
            P, P, P, P, P, D, I, P, P, P, U, P, I, I, D, P,  # step part
            P, P, P, P, I, I, P,  # rel_in, rel_out, anchor, ustep; orthorhombic; streamed words
        ],
    ),
    "tsx_md_reduce": (I, [P, P, I, I, P, P, I, P, P, P]),
    "tsx_copy_ref": (I, [P, P, L, P, P, P]),
    "tsx_npt_ws_bytes": (L, [L, I]),
    "tsx_npt_langevin_iso_pre": (
        I,
        [I, P, P, L, P, P, P, P, P, P, P, P, P, D, P, P, P, P, P, P, P, P, P, P],
    ),
    "tsx_npt_langevin_iso_post": (
        I,
        [I, P, P, L, P, P, P, P, P, P, P, P, D, P, P, P, P, P, P, P],
    ),
    "tsx_npt_langevin_aniso_pre": (
        I,
        [I, P, P, L, P, P, P, P, P, P, P, P, P, D, P, P, P, P, P, P, P, P, P, P],
    ),
    "tsx_npt_langevin_aniso_post": (
        I,
        [I, P, P, L, P, P, P, P, P, P, P, P, D, P, P, P, P, P, P],
    ),
    "tsx_mom_reduce_ws_bytes": (L, [I, I]),
    "tsx_md_scale_kick_reduce": (I, [I, P, I, P, P, P, P, D, I, D, P, P, P]),
    "tsx_md_scale_kick_drift": (I, [L, P, P, P, P, P, D, P, D, P]),
    "tsx_nhc_half_step": (I, [I, I, I, I, D, P, P, P, P, P, P, P, I, P, P, P]),
    "tsx_vrescale_factor": (I, [I, P, P, P, P, D, P, P, P, P]),
    "tsx_npt_nh_ws_bytes": (L, [I]),
    "tsx_npt_nh_pre": (
        I,
        [I, P, P, L, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, D, P, P],
    ),
    "tsx_npt_nh_post": (
        I,
        [I, P, I, P, P, P, P, P, P, P, P, P, D, P, P, P, P, P, P],
    ),
    "tsx_crescale_ws_bytes": (L, [I]),
    "tsx_crescale_step": (
        I,
        [I, I, P, L, P, P, P, P, P, P, P, P, P, D, P, P, P, P, P, P, P, P],
    ),
    "tsx_fire_ws_bytes": (L, [I]),
    "tsx_fire_pre": (
        I,
        [P, I, P, P, L, P, I, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, D, D, P, P, P],
    ),
    "tsx_grad_ws_bytes": (L, [L, I, I]),
    "tsx_pair_grad": (
        I,
        [I, P, I, P, P, I, P, P, P, I, P, P, P, L, I, I, L, D, P, P, P, P, P, I, P, P],
    ),
    "tsx_nve_adjoint_update": (I, [P, P, P, P, D, D, P, L, P]),
    "tsx_fire_post": (
        I,
        [P, I, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, I, P],
    ),
}


class FireParams(ctypes.Structure):
    """tsx_fire_params (include/tsx.h)."""

    _fields_ = [
        ("dt_max", ctypes.c_double),
        ("f_inc", ctypes.c_double),
        ("f_dec", ctypes.c_double),
        ("alpha_start", ctypes.c_double),
        ("f_alpha", ctypes.c_double),
        ("max_step", ctypes.c_double),
        ("eps", ctypes.c_double),
        ("n_min", ctypes.c_int32),
        ("flavor", ctypes.c_int32),
        ("cell_filter", ctypes.c_int32),
        ("hydrostatic_strain", ctypes.c_int32),
        ("constant_volume", ctypes.c_int32),
        ("first_step", ctypes.c_int32),
    ]


class NLFrame(ctypes.Structure):
    """tsx_nl_frame (include/tsx.h): bin-relative frame of a stencil-word list."""

    _fields_ = [
        ("anchor", ctypes.c_void_p),
        ("rel", ctypes.c_void_p),
        ("ustep", ctypes.c_void_p),
        ("e_atom", ctypes.c_void_p),
        ("energy", ctypes.c_void_p),
        ("gate_clear", ctypes.c_void_p),
        ("step", ctypes.c_void_p),
    ]


_lib = None


class TsxError(RuntimeError):
    """Raised when a tsx C-ABI call returns a non-zero status."""


def lib() -> ctypes.CDLL:
    """Load libtsx.so once (after torch) and declare every signature."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise TsxError(
                f"{_LIB_PATH} is missing: build it with `python torch-sim_amd/_build.py`"
            )
        cdll = ctypes.CDLL(str(_LIB_PATH), mode=ctypes.RTLD_GLOBAL)
        for name, (res, args) in _SIGS.items():
            fn = getattr(cdll, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = cdll
    return _lib


def exported_symbols() -> list[str]:
    """Names from the signature table that the loaded library exports."""
    cdll = lib()
    return [n for n in _SIGS if hasattr(cdll, n)]


def call(name: str, *args) -> int:
    """Invoke a tsx entry point; raise TsxError with tsx_last_error() on failure."""
    fn = getattr(lib(), name)
    rc = fn(*args)
    if isinstance(rc, int) and rc != 0 and name not in ("tsx_version",):
        msg = lib().tsx_last_error().decode(errors="replace")
        raise TsxError(f"{name} failed (rc={rc}): {msg}")
    return rc


def ptr(t: torch.Tensor | None) -> int | None:
    """Device pointer of a tensor (None for NULL)."""
    if t is None:
        return None
    return t.data_ptr()


def stream() -> int:
    """hipStream_t of torch's current stream."""
    return torch.cuda.current_stream().cuda_stream


def require_gpu(t: torch.Tensor, what: str = "tensor") -> None:
    """The product path runs only on a ROCm GPU; refuse CPU tensors loudly."""
    if t.device.type != "cuda":
        raise TsxError(
            f"torch_sim_amd runs its kernels on a ROCm GPU; {what} is on {t.device}"
        )
