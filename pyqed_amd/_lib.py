"""ctypes binding of libpyqed_amd.so (declared in include/pyqed_amd.h).

The library is loaded from this package directory.  There is no CPU fallback:
when the library or a GPU is missing, the propagators raise instead of
computing anything another way.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# PYQED_AMD_LIB selects another build of the library (e.g. a variant compiled with other tuning).
LIB_PATH = os.environ.get("PYQED_AMD_LIB") or os.path.join(_HERE, "libpyqed_amd.so")

PYQED_OK = 0
PREC_FP64 = 0
PREC_MIXED = 1
PLAN_NO_RESIDENT = 1
PLAN_DENSE_Q = 2
PLAN_GENERIC = 4
PLAN_HERMITIAN = 8
PLAN_PACKED = 16
PLAN_ADJOINT = 32

SYMBOLS = [
    "pyqed_last_error", "pyqed_version", "pyqed_device_count",
    "pyqed_heom_plan_create", "pyqed_heom_plan_destroy", "pyqed_heom_plan_query", "pyqed_heom_plan_layout",
    "pyqed_heom_tables_export", "pyqed_heom_tables_size", "pyqed_heom_tables_build", "pyqed_heom_run", "pyqed_heom_rhs", "pyqed_heom_rhs_combine", "pyqed_heom_plan_set_class_tiling", "pyqed_heom_plan_get_class_tiling", "pyqed_lindblad_work_bytes", "pyqed_lindblad_run", "pyqed_heom_plan_stats",
    "pyqed_heom_plan_classes", "pyqed_heom_plan_class_single", "pyqed_heom_plan_set_tiling", "pyqed_heom_plan_get_tiling", "pyqed_heom_plan_set_variant",
    "pyqed_heom_plan_class_state", "pyqed_heom_plan_set_timing", "pyqed_heom_plan_read_timing",
    "pyqed_heom_adjoint_work_bytes", "pyqed_heom_adjoint_contract",
    "pyqed_liouville_run", "pyqed_liouville_adjoint", "pyqed_legacy_heom_run",
    "pyqed_spectral_work_bytes", "pyqed_spectral_sum", "pyqed_spectral_sum_static",
]
SPECTRAL_STATIC_MAX = 1 << 20  # complex weights the static scratch of pyqed_spectral_sum_static holds


class PyqedError(RuntimeError):
    """A call into libpyqed_amd failed (the message is the library's own)."""


class HeomInstance(C.Structure):
    _fields_ = [
        ("d", C.c_int32), ("K", C.c_int32), ("L", C.c_int32), ("ncoup", C.c_int32),
        ("mode", C.c_void_p), ("expn", C.c_void_p), ("etal", C.c_void_p), ("etar", C.c_void_p),
        ("etaa", C.c_void_p), ("H", C.c_void_p), ("sdip", C.c_void_p), ("qdiag", C.c_void_p),
        ("cdiag", C.c_void_p), ("qmat", C.c_void_p), ("cmat", C.c_void_p),
    ]


_lib = None


def lib():
    """Load libpyqed_amd.so (once) and declare argument types."""
    global _lib
    if _lib is not None:
        return _lib
    # torch first: its bundled HIP runtime then also serves this library (same soname), so torch
    # streams and allocations are valid handles here.
    import torch  # noqa: F401
    if not os.path.exists(LIB_PATH):
        raise PyqedError(f"{LIB_PATH} is missing: build it with `python -m pyqed_amd.build` "
                         "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.pyqed_last_error.restype = C.c_char_p
    L.pyqed_last_error.argtypes = []
    L.pyqed_version.restype = C.c_int
    L.pyqed_device_count.argtypes = [C.POINTER(C.c_int)]
    L.pyqed_heom_plan_create.argtypes = [C.POINTER(HeomInstance), i32, i32, C.POINTER(vp)]
    L.pyqed_heom_plan_destroy.argtypes = [vp]
    L.pyqed_heom_plan_query.argtypes = [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]
    L.pyqed_heom_plan_layout.argtypes = [vp, vp, vp, vp]
    L.pyqed_heom_tables_export.argtypes = [vp, i32, vp, vp, vp]
    L.pyqed_heom_tables_size.argtypes = [i32, i32, C.POINTER(i64)]
    L.pyqed_heom_tables_build.argtypes = [i32, i32, vp, vp, vp]
    L.pyqed_heom_run.argtypes = [vp, i32, vp, vp, i64, vp, i64, i64, dbl, vp, i64, vp]
    L.pyqed_heom_rhs.argtypes = [vp, vp, vp, vp, i64, i32, vp]
    L.pyqed_heom_rhs_combine.argtypes = [vp, vp, vp, vp, C.c_double, vp, C.c_double, i32, vp, i64, i32, vp]
    L.pyqed_lindblad_work_bytes.argtypes = [i32, i32, i32, i32, C.POINTER(i64)]
    L.pyqed_heom_plan_set_class_tiling.argtypes = [vp, i32, i64, i32]
    L.pyqed_heom_plan_get_class_tiling.argtypes = [vp, vp, vp]
    L.pyqed_lindblad_run.argtypes = [i32, i32, vp, vp, vp, vp, i32, vp, vp, i32, vp, vp, i64, C.c_double, i32, vp, vp,
                                     vp, i64, vp, i64, vp]
    L.pyqed_heom_plan_stats.argtypes = [vp, vp]
    L.pyqed_heom_plan_classes.argtypes = [vp, vp, vp]
    L.pyqed_heom_plan_class_single.argtypes = [vp, vp]
    L.pyqed_heom_plan_set_tiling.argtypes = [vp, i64, i32]
    L.pyqed_heom_plan_get_tiling.argtypes = [vp, C.POINTER(i64), C.POINTER(i32)]
    L.pyqed_heom_plan_set_variant.argtypes = [vp, i32]
    L.pyqed_heom_plan_class_state.argtypes = [vp, vp]
    L.pyqed_heom_plan_set_timing.argtypes = [vp, i32]
    L.pyqed_heom_plan_read_timing.argtypes = [vp, vp, vp]
    L.pyqed_heom_adjoint_work_bytes.argtypes = [vp, C.POINTER(i64)]
    L.pyqed_heom_adjoint_contract.argtypes = [vp, vp, vp, vp, vp, vp, i64, vp]
    L.pyqed_liouville_run.argtypes = [i32, i32, vp, i32, vp, vp, i64, dbl, i32, vp, vp, vp, i64, vp]
    L.pyqed_liouville_adjoint.argtypes = [i32, i32, vp, i32, vp, vp, i64, dbl, i32, vp, vp, vp, vp, vp, i64, vp]
    L.pyqed_legacy_heom_run.argtypes = [i32, i32, i32, i32, vp, vp, vp, vp, vp, i64, dbl, i32, vp, vp, vp]
    L.pyqed_spectral_work_bytes.argtypes = [i32, i32]
    L.pyqed_spectral_sum.argtypes = [i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, i64, i64, i64, vp, i64, vp]
    L.pyqed_spectral_sum_static.argtypes = [i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, i64, i64, i64, vp]
    for name in SYMBOLS:
        if name != "pyqed_last_error":
            getattr(L, name).restype = C.c_int
    L.pyqed_spectral_work_bytes.restype = C.c_int64
    _lib = L
    return L


def check(rc):
    if rc != PYQED_OK:
        raise PyqedError(f"libpyqed_amd error {rc}: {lib().pyqed_last_error().decode()}")


def require_gpu():
    """Return torch with a usable GPU, or raise (the propagators have no CPU path)."""
    import torch
    lib()
    if not torch.cuda.is_available():
        raise PyqedError("pyqed_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch


def ptr(t):
    """Device (or host) address of a tensor / numpy array, or None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def stream_ptr(torch, stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def hierarchy_tables(K, L):
    """keys [nmax, K] int64, plus / minus [nmax, K] int32 built by the library's C++ (no GPU needed)."""
    import numpy as np
    lb = lib()
    n = C.c_int64()
    check(lb.pyqed_heom_tables_size(K, L, C.byref(n)))
    keys = np.zeros((n.value, K), np.int64)
    plus = np.zeros((n.value, K), np.int32)
    minus = np.zeros((n.value, K), np.int32)
    check(lb.pyqed_heom_tables_build(K, L, keys.ctypes.data, plus.ctypes.data, minus.ctypes.data))
    return keys, plus, minus
