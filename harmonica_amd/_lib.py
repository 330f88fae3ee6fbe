"""
ctypes binding of libhmx.so, the C-ABI declared in include/hmx.h.

torch is imported before the library is loaded so that libhmx.so binds to
the HIP runtime torch already loaded (both carry the soname
libamdhip64.so.7); device pointers of torch tensors are then valid in the
library.  There is no CPU fallback: if the library or a GPU is missing,
calls raise.
"""

import ctypes
import os

import torch  # noqa: F401  (must precede the dlopen of libhmx.so)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhmx.so")

HMX_OK = 0
HMX_ERR_ARG = -1
HMX_ERR_HIP = -2
HMX_ERR_NODEV = -3
HMX_ERR_NUMERIC = -4

HMX_F64 = 0
HMX_F32 = 1
HMX_CARTESIAN = 0
HMX_SPHERICAL = 1
HMX_FIT_FP64 = 0
HMX_FIT_FP32_MFMA = 1
HMX_FIT_BF16_SPLIT = 2
HMX_FIT_INT8_SPLIT = 3
HMX_FIT_NO_DAMPING = 0x100
HMX_SVD_RIDGE = 0
HMX_SVD_LSTSQ = 1
HMX_SVD_DUAL_LSTSQ = 2

#: Field names in the reference's FIELDS order (gravity.py:37-48); bit i of a
#: field mask selects FIELDS[i].
FIELDS = ("potential", "g_e", "g_n", "g_z", "g_ee", "g_nn", "g_zz", "g_en", "g_ez", "g_nz")
FIELD_BIT = {f: 1 << i for i, f in enumerate(FIELDS)}

# Every symbol exported by include/hmx.h, with (restype, argtypes).
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32
_int = ctypes.c_int
_dbl = ctypes.c_double
_szp = ctypes.POINTER(ctypes.c_size_t)
SIGNATURES = {
    "hmx_last_error": (ctypes.c_char_p, []),
    "hmx_version": (_int, []),
    "hmx_device_count": (_int, [ctypes.POINTER(ctypes.c_int)]),
    "hmx_forward_workspace_bytes": (_int, [_i64, _i64, _u32, _szp]),
    "hmx_prism_gravity": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _u32, _vp, _vp, _vp,
                                 ctypes.c_size_t, _vp]),
    "hmx_prism_gravity_static": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _u32, _vp, _vp,
                                        _vp]),
    "hmx_prism_singular_scan": (_int, [_vp, _vp, _vp, _i64, _vp, _i64, _u32, _vp, _vp]),
    "hmx_prism_layer_gravity": (_int, [_vp, _vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _vp,
                                       _dbl, _u32, _vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "hmx_prism_magnetic_workspace_bytes": (_int, [_i64, _i64, _u32, _szp]),
    "hmx_prism_magnetic": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _i64, _u32, _vp, _vp,
                                  ctypes.c_size_t, _vp]),
    "hmx_point_gravity": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _i64, _int, _u32, _vp,
                                 _vp, ctypes.c_size_t, _vp]),
    "hmx_eqs_jacobian": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _int, _vp, _i64, _vp]),
    "hmx_eqs_fit_workspace_bytes": (_int, [_i64, _i64, _int, _szp]),
    "hmx_eqs_fit": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _dbl, _int, _int,
                           _dbl, _vp, ctypes.POINTER(ctypes.c_double), _vp, ctypes.c_size_t, _vp]),
    "hmx_lsq_svd_workspace_bytes": (_int, [_i64, _i64, _szp]),
    "hmx_lsq_svd": (_int, [_vp, _i64, _i64, _vp, _dbl, _int, _dbl, _vp, _vp,
                           ctypes.POINTER(ctypes.c_double), _vp, ctypes.c_size_t, _vp]),
    "hmx_tesseroid_workspace_bytes": (_int, [_i64, _i64, _szp]),
    "hmx_tesseroid_gravity": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _u32, _int, _vp,
                                     _vp, _vp, ctypes.c_size_t, _vp]),
    "hmx_geocentric": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp]),
    "hmx_eqs_gb_fit_workspace_bytes": (_int, [_i64, _i64, _i64, _int, _dbl, _i64, _szp]),
    "hmx_eqs_gb_fit": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                              _vp, _i64, _dbl, _int, _i64, _int, _dbl, _vp, _vp, _vp,
                              ctypes.POINTER(ctypes.c_double), _vp, ctypes.c_size_t, _vp]),
    "hmx_gz_grad_workspace_bytes": (_int, [_i64, _i64, _szp]),
    "hmx_prism_gz_jvp": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, ctypes.c_size_t,
                                _vp]),
    "hmx_prism_gz_vjp": (_int, [_vp, _vp, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp, ctypes.c_size_t,
                                _vp]),
    "hmx_prism_layer_gz_jvp": (_int, [_vp, _vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _dbl,
                                      _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "hmx_prism_layer_gz_vjp": (_int, [_vp, _vp, _vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _vp,
                                      _dbl, _vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "hmx_batched_plan_size": (_int, [_vp, _vp, _i64, _i64, ctypes.POINTER(ctypes.c_int64)]),
    "hmx_batched_plan": (_int, [_vp, _vp, _i64, _i64, _vp, _i64, ctypes.POINTER(ctypes.c_int64), _vp,
                                _vp]),
    "hmx_batched_workspace_bytes": (_int, [_i64, _szp]),
    "hmx_batched_prism_gz": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _i64, _i64,
                                    _int, _vp, _vp, ctypes.c_size_t, _vp]),
    "hmx_mfma_probe": (_int, [_int, _vp, _vp, _vp, _vp]),
    "hmx_gemm_tn": (_int, [_int, _vp, _i64, _vp, _i64, _i64, _i64, _i64, _vp, _i64, _dbl, _dbl,
                           _dbl, _int, _vp]),
    "hmx_cholesky_workspace_bytes": (_int, [_i64, _szp]),
    "hmx_cholesky_solve": (_int, [_vp, _i64, _vp, ctypes.POINTER(ctypes.c_int), _vp,
                                  ctypes.c_size_t, _vp]),
    "hmx_devmath_eval": (_int, [_int, _vp, _vp, _vp, _i64, _vp]),
    "hmx_eqs_predict_workspace_bytes": (_int, [_i64, _i64, _szp]),
    "hmx_eqs_predict": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _i64, _int, _vp, _vp,
                               ctypes.c_size_t, _vp]),
}

_lib = None


class HMXError(RuntimeError):
    """A libhmx call returned a non-zero status."""


def lib():
    """Load libhmx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m harmonica_amd.csrc.build` "
                "(harmonica_amd has no CPU fallback)"
            )
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    """Raise HMXError with the library's message on a non-zero status."""
    if status != HMX_OK:
        msg = lib().hmx_last_error().decode(errors="replace")
        raise HMXError(f"libhmx error {status}: {msg}")


def require_device():
    """Return the torch device the kernels run on; raise if no gfx950 GPU."""
    if not torch.cuda.is_available():
        raise RuntimeError("harmonica_amd needs a ROCm GPU (MI355X/gfx950); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    """Device (or host) address of a tensor, or None for None."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def field_mask(fields):
    m = 0
    for f in fields:
        m |= FIELD_BIT[f]
    return m
