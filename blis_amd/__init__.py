"""Python host helpers for the MI355X BLIS library.

The product is the C library ``blis_amd/lib/libblis_amd.so`` (declared in
``include/blis.h`` and ``include/blis_amd.h``). This module only loads it with
ctypes and provides thin, typed wrappers with BLIS argument meaning, for the
tests, the smoke check and the benchmark. Device buffers are plain pointers
(for example ``torch.Tensor.data_ptr()`` of a HIP tensor).

There is no fallback: if the library is missing, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libblis_amd.so"

# Enum encodings (frame/include/bli_type_defs.h:254-520).
BLIS_NO_TRANSPOSE = 0x0
BLIS_TRANSPOSE = 0x8
BLIS_CONJ_NO_TRANSPOSE = 0x10
BLIS_CONJ_TRANSPOSE = 0x18
BLIS_NO_CONJUGATE = 0x0
BLIS_CONJUGATE = 0x10
BLIS_LOWER = 0xC0
BLIS_UPPER = 0x60
BLIS_DENSE = 0xE0
BLIS_LEFT = 0
BLIS_RIGHT = 1
BLIS_NONUNIT_DIAG = 0x0
BLIS_UNIT_DIAG = 0x100
BLIS_FLOAT = 0
BLIS_DOUBLE = 2
BLIS_SYMMETRIC = 0x2 << 20
BLIS_TRIANGULAR = 0x3 << 20

dim_t = C.c_int64
inc_t = C.c_int64
f77_int = C.c_int32


class atom_t(C.Structure):
    _fields_ = [("real", C.c_double), ("imag", C.c_double)]


class obj_t(C.Structure):
    """Layout of BLIS's obj_t (frame/include/bli_type_defs.h:1233-1263)."""


obj_t._fields_ = [
    ("root", C.POINTER(obj_t)),
    ("off", dim_t * 2),
    ("dim", dim_t * 2),
    ("diag_off", dim_t),
    ("info", C.c_uint32),
    ("info2", C.c_uint32),
    ("elem_size", C.c_uint64),
    ("buffer", C.c_void_p),
    ("rs", inc_t),
    ("cs", inc_t),
    ("is_", inc_t),
    ("scalar", atom_t),
    ("m_padded", dim_t),
    ("n_padded", dim_t),
    ("ps", inc_t),
    ("pd", inc_t),
    ("m_panel", dim_t),
    ("n_panel", dim_t),
]

_lib = None

P = C.c_void_p
D = C.c_double
F = C.c_float


def _sig(lib):
    """Attach argtypes for the entry points the Python side calls."""
    def s(name, res, *args):
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = list(args)

    for ch, T in (("d", D), ("s", F)):
        PT = C.POINTER(T)
        s(f"bli_{ch}gemm", None, C.c_int, C.c_int, dim_t, dim_t, dim_t, PT, P, inc_t, inc_t,
          P, inc_t, inc_t, PT, P, inc_t, inc_t)
        s(f"bli_{ch}trsm", None, C.c_int, C.c_int, C.c_int, C.c_int, dim_t, dim_t, PT, P,
          inc_t, inc_t, P, inc_t, inc_t)
        s(f"bli_{ch}axpyv", None, C.c_int, dim_t, PT, P, inc_t, P, inc_t)
        s(f"bli_{ch}dotv", None, C.c_int, C.c_int, dim_t, P, inc_t, P, inc_t, P)
        s(f"bli_{ch}gemv", None, C.c_int, C.c_int, dim_t, dim_t, PT, P, inc_t, inc_t, P, inc_t,
          PT, P, inc_t)
        s(f"bli_{ch}gemm_batch", None, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, dim_t, P)
        s(f"bli_amd_{ch}gemm_batch_plan_create", P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P,
          P, dim_t, P)
        PF = C.POINTER(f77_int)
        PC = C.c_char_p
        s(f"{ch}gemm_", None, PC, PC, PF, PF, PF, PT, P, PF, P, PF, PT, P, PF)
        s(f"{ch}trsm_", None, PC, PC, PC, PC, PF, PF, PT, P, PF, P, PF)
        s(f"{ch}gemv_", None, PC, PF, PF, PT, P, PF, P, PF, PT, P, PF)
        s(f"{ch}axpy_", None, PF, PT, P, PF, P, PF)
        s(f"{ch}dot_", T, PF, P, PF, P, PF)
        s(f"{ch}gemm_batch_", None, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P)
        # structured level-3 (typed and BLAS)
        s(f"bli_{ch}gemmt", None, C.c_int, C.c_int, C.c_int, dim_t, dim_t, PT, P, inc_t, inc_t,
          P, inc_t, inc_t, PT, P, inc_t, inc_t)
        s(f"bli_{ch}syrk", None, C.c_int, C.c_int, dim_t, dim_t, PT, P, inc_t, inc_t, PT, P,
          inc_t, inc_t)
        s(f"bli_{ch}syr2k", None, C.c_int, C.c_int, C.c_int, dim_t, dim_t, PT, P, inc_t, inc_t,
          P, inc_t, inc_t, PT, P, inc_t, inc_t)
        s(f"bli_{ch}symm", None, C.c_int, C.c_int, C.c_int, C.c_int, dim_t, dim_t, PT, P, inc_t,
          inc_t, P, inc_t, inc_t, PT, P, inc_t, inc_t)
        s(f"bli_{ch}trmm", None, C.c_int, C.c_int, C.c_int, C.c_int, dim_t, dim_t, PT, P, inc_t,
          inc_t, P, inc_t, inc_t)
        s(f"bli_{ch}trmm3", None, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dim_t, dim_t, PT,
          P, inc_t, inc_t, P, inc_t, inc_t, PT, P, inc_t, inc_t)
        s(f"{ch}symm_", None, PC, PC, PF, PF, PT, P, PF, P, PF, PT, P, PF)
        s(f"{ch}trmm_", None, PC, PC, PC, PC, PF, PF, PT, P, PF, P, PF)
        s(f"{ch}syrk_", None, PC, PC, PF, PF, PT, P, PF, PT, P, PF)
        s(f"{ch}syr2k_", None, PC, PC, PF, PF, PT, P, PF, P, PF, PT, P, PF)
        s(f"{ch}gemmt_", None, PC, PC, PC, PF, PF, PT, P, PF, P, PF, PT, P, PF)
        s(f"bli_amd_{ch}gemm_split", C.c_int, C.c_int, C.c_int, dim_t, dim_t, dim_t, PT, P,
          inc_t, inc_t, P, inc_t, inc_t, PT, P, inc_t, inc_t, C.c_int, C.POINTER(C.c_double))
    for ch in ("z", "c"):
        s(f"bli_{ch}gemm", None, C.c_int, C.c_int, dim_t, dim_t, dim_t, P, P, inc_t, inc_t, P,
          inc_t, inc_t, P, P, inc_t, inc_t)
    s("bli_amd_gemm_noalloc", C.c_int, C.c_int, C.c_int, C.c_int, dim_t, dim_t, dim_t, P, P,
      inc_t, inc_t, P, inc_t, inc_t, P, P, inc_t, inc_t, P)
    s("bli_amd_gemm_batch_plan_execute", None, P, C.c_int)
    s("bli_amd_gemm_batch_plan_destroy", None, P)
    s("bli_amd_gemm_batch_plan_work_items", C.c_int, P)
    s("bli_amd_gemm_batch_plan_autotune", C.c_int, P, C.c_int)
    s("bli_amd_gemm_batch_plan_set_mode", None, P, C.c_int)
    s("bli_amd_gemm_batch_plan_get_mode", C.c_int, P)
    s("bli_amd_gemm_batch_plan_set_single_dispatch", None, P, C.c_int)
    s("bli_amd_gemm_batch_plan_get_single_dispatch", C.c_int, P)
    s("bli_amd_set_stream", None, P, C.c_int)
    s("bli_amd_get_stream", P)
    s("bli_amd_synchronize", None)
    s("bli_amd_gemm_split_slices", C.c_int, C.c_int, C.c_double)
    s("bli_amd_gemm_split_max_slices", C.c_int, C.c_int)
    s("bli_info_get_version_str", C.c_char_p)
    s("bli_gemm", None, P, P, P, P, P)
    s("bli_trsm", None, C.c_int, P, P, P)
    s("bli_gemmt", None, P, P, P, P, P)
    s("bli_syrk", None, P, P, P, P)
    s("bli_syr2k", None, P, P, P, P, P)
    s("bli_symm", None, C.c_int, P, P, P, P, P)
    s("bli_trmm", None, C.c_int, P, P, P)
    s("bli_trmm3", None, C.c_int, P, P, P, P, P)
    s("bli_obj_create_with_attached_buffer", None, C.c_int, dim_t, dim_t, P, inc_t, inc_t, P)
    s("bli_obj_create_1x1_with_attached_buffer", None, C.c_int, P, P)
    if hasattr(lib, "bli_amd_gemm_set_config"):
        s("bli_amd_gemm_set_config", None, C.c_int, C.c_int, C.c_int, C.c_int)


def lib():
    """Load libblis_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing; build it with `python -c "
                f"'import __graft_entry__ as g; g.build()'` or `python blis_amd/build.py`")
        # torch ships its own HIP runtime (libamdhip64 / libhsa-runtime64).
        # Load it first so that this library binds to the same runtime
        # instance instead of bringing up a second one in the process.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = C.CDLL(str(LIB_PATH), mode=C.RTLD_LOCAL)
        _sig(_lib)
    return _lib


def ptr(x) -> int:
    """Device (or host) address of a torch tensor or numpy array."""
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    return x.ctypes.data


def build(verbose: bool = False):
    from .build import build as _build
    return _build(verbose=verbose)
