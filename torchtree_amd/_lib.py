"""ctypes binding of ``libhiptree.so`` (C ABI declared in ``include/hiptree.h``).

The library is loaded after ``torch`` so that both share the one HIP runtime
(``libamdhip64.so.7``) already mapped by torch; a second runtime in the
process is detected and refused. There is no CPU fallback: every entry
point raises ``HipTreeError`` when the library is missing or a call fails.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch  # noqa: F401  (must be imported before the library is mapped)

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libhiptree.so"

HX_F32 = 0
HX_F64 = 1
HX_F32_BF16X3 = 2  # FP32 partials, dense-state products as split BF16
HX_TIPS_CODES = 0
HX_TIPS_DENSE = 1

_c_int = ctypes.c_int
_c_size = ctypes.c_size_t
_vp = ctypes.c_void_p
_dp = ctypes.c_void_p  # device or host double* passed as raw address

# name -> (restype, argtypes)
_SIGNATURES = {
    "hx_error_string": (ctypes.c_char_p, [_c_int]),
    "hx_last_error": (ctypes.c_char_p, []),
    "hx_version": (_c_int, []),
    "hx_plan_create": (_c_int, [_c_int, _vp, _c_int, ctypes.POINTER(_vp)]),
    "hx_plan_destroy": (_c_int, [_vp]),
    "hx_plan_info": (_c_int, [_vp, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int),
                              ctypes.POINTER(_c_int)]),
    "hx_plan_walk_info": (_c_int, [_vp] + [ctypes.POINTER(_c_int)] * 5),
    "hx_workspace_bytes": (_c_size, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int]),
    "hx_loglik": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int,
                           _c_int, _vp, _vp, _c_int,
                           _dp, _vp, _vp, _c_int, _vp, _c_int,
                           _vp, _c_size, _dp, _dp, _dp, _vp]),
    "hx_loglik_grad_mats": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int,
                                     _c_int, _vp, _vp, _c_int,
                                     _dp, _vp, _vp, _c_int, _vp, _c_int,
                                     _vp, _c_size, _dp, _vp]),
    "hx_loglik_grad_tau": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int,
                                    _c_int, _vp, _vp, _c_int,
                                    _dp, _vp, _vp, _vp, _c_int, _vp, _c_int,
                                    _vp, _c_size, _dp, _vp]),
    "hx_pmatrix": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int,
                            _dp, _dp, _dp, _dp, _dp, _c_int, _vp, _c_int, _vp]),
    "hx_pmatrix_dtau": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int,
                                 _dp, _dp, _dp, _dp, _dp, _c_int, _vp, _c_int, _vp]),
    "hx_pmatrix_grad_scratch_bytes": (_c_size, [_c_int, _c_int, _c_int, _c_int]),
    "hx_pmatrix_grad": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int,
                                 _dp, _dp, _dp, _dp, _dp, _c_int,
                                 _dp, _dp, _dp, _dp, _vp, _vp]),
    "hx_pmatrix_expm": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int,
                                 _dp, _dp, _dp, _c_int, _vp, _c_int, _vp]),
    "hx_pmatrix_expm_dtau": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int,
                                      _dp, _dp, _dp, _c_int, _vp, _c_int, _vp]),
    "hx_pmatrix_expm_grad_scratch_bytes": (_c_size, [_c_int, _c_int, _c_int, _c_int]),
    "hx_pmatrix_expm_grad": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int,
                                      _dp, _dp, _dp, _c_int, _dp, _dp, _dp, _dp, _vp, _vp]),
    "hx_job_create": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int,
                               _vp, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int,
                               ctypes.POINTER(_vp)]),
    "hx_job_create_general": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int,
                                       _vp, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _c_int,
                                       ctypes.POINTER(_vp)]),
    "hx_job_destroy": (_c_int, [_vp]),
    "hx_job_resident_bytes": (_c_size, [_vp]),
    "hx_job_workspace_bytes": (_c_size, [_vp, _c_int, _c_int]),
    "hx_job_workspace_bytes_for": (_c_size, [_c_int] * 7),
    "hx_job_eval": (_c_int, [_vp, _c_int, _dp, _dp, _dp, _vp, _c_size, _vp]),
    "hx_job_eval_noalloc": (_c_int, [_vp, _c_int, _dp, _dp, _dp, _vp]),
    "hx_job_model_workspace_bytes": (_c_size, [_vp, _c_int, _c_int]),
    "hx_job_eval_model": (_c_int, [_vp, _c_int, _dp, _c_int, _dp, _dp, _dp, _dp, _dp, _dp, _c_int,
                                   _dp, _dp, _dp, _dp, _dp, _dp, _vp, _c_size, _vp]),
    "hx_selftest_mfma_layout": (_c_int, [ctypes.POINTER(_c_int), _vp]),
}

HX_NOALLOC_WORKSPACE_BYTES = 64 << 20  # include/hiptree.h

HX_GRAD_NONE = 0
HX_GRAD_MATS = 1
HX_GRAD_TAU = 2

EXPORTED = tuple(_SIGNATURES)

_lib = None


class HipTreeError(RuntimeError):
    pass


def _mapped_hip_runtimes():
    paths = set()
    try:
        with open("/proc/self/maps") as fp:
            for line in fp:
                if "libamdhip64.so" in line:
                    paths.add(os.path.realpath(line.split()[-1]))
    except OSError:
        pass
    return paths


def load():
    """Load and declare the library (idempotent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise HipTreeError(
            f"{LIB_PATH} is missing: build it with `python -m torchtree_amd.build`")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    runtimes = _mapped_hip_runtimes()
    if len(runtimes) > 1:
        raise HipTreeError(f"two HIP runtimes mapped in this process: {sorted(runtimes)}")
    _lib = lib
    return lib


def check(status: int, what: str = "") -> None:
    if status != 0:
        lib = load()
        msg = lib.hx_error_string(status).decode()
        detail = lib.hx_last_error().decode()
        raise HipTreeError(f"{what}: {msg} ({detail})" if what else f"{msg} ({detail})")


def ptr(t) -> int:
    """Raw address of a tensor's storage start (None -> 0)."""
    if t is None:
        return 0
    return t.data_ptr()


def stream_handle(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def require_gpu(t: torch.Tensor, what: str) -> None:
    if not t.is_cuda:
        raise HipTreeError(f"{what} must be a device tensor (HIP engine, no CPU path)")
