"""ctypes binding of libaclblas_cgemm3m.so (the C ABI declared in include/aclblas_cgemm3m.h).

The library is the only compute path: if it is missing this module raises, there is no
fallback. Build it with ``python ascend-open-source-contributions_amd/build.py``.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libaclblas_cgemm3m.so"))

# ops-blas/docs/zh/api_list.md:31-42
STATUS = {
    0: "SUCCESS", 1: "NOT_INITIALIZED", 2: "ALLOC_FAILED", 3: "INVALID_VALUE", 4: "MAPPING_ERROR",
    5: "EXECUTION_FAILED", 6: "INTERNAL_ERROR", 7: "NOT_SUPPORTED", 8: "ARCH_MISMATCH",
    9: "HANDLE_IS_NULLPTR", 10: "INVALID_ENUM", 11: "UNKNOWN",
}
SUCCESS, INVALID_VALUE, EXECUTION_FAILED, NOT_SUPPORTED, HANDLE_IS_NULLPTR, INVALID_ENUM = 0, 3, 5, 7, 9, 10
OP = {"N": 111, "T": 112, "C": 113}
PRECISION = {"fp32": 0, "bf16x3": 1, "fp16": 2, "bf16x6": 3}

# Every symbol include/aclblas_cgemm3m.h declares.
EXPORTS = [
    "aclblasCreate", "aclblasDestroy", "aclblasSetStream", "aclblasGetStream", "aclblasSetWorkspace",
    "aclblasGetVersion", "aclblasCgemm3m", "aclblasCgemm", "aclblasCgemm3mStridedBatched", "aclblasCgemm3mEx",
    "aclblasCgemm3mGetWorkspaceSize",
    "aclblasCgemm3mGroupedBatched", "aclblasCgemm3mGroupedBatchedEx", "aclblasCgemm3mGroupedBatchedGetWorkspaceSize",
    "aclblasCgemm3mResidual", "aclblasCgemm3mGetLastPlan", "aclblasCgemm3mSetTiming",
    "aclblasCgemm3mGetTiming", "aclblasSetLogCallback", "aclblasCgemm3mNoAlloc",
]
STAGES = ["pack", "energy", "gemm", "merge", "recompute", "scale"]


class Complex(ctypes.Structure):
    _fields_ = [("real", ctypes.c_float), ("imag", ctypes.c_float)]


class Config(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("tile", ctypes.c_int), ("kSplit", ctypes.c_int),
                ("correction", ctypes.c_int), ("identity", ctypes.c_int), ("variant", ctypes.c_int)]


LOG_CB = ctypes.CFUNCTYPE(None, ctypes.c_char_p)
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libaclblas_cgemm3m.so is not built (%s); run build.py" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, i, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
    cp = ctypes.POINTER(Complex)
    L.aclblasCreate.argtypes = [ctypes.POINTER(vp)]
    L.aclblasDestroy.argtypes = [vp]
    L.aclblasSetStream.argtypes = [vp, vp]
    L.aclblasGetStream.argtypes = [vp, ctypes.POINTER(vp)]
    L.aclblasSetWorkspace.argtypes = [vp, vp, sz]
    L.aclblasGetVersion.argtypes = [vp, ctypes.POINTER(i)]
    gemm = [vp, i, i, i, i, i, cp, vp, i, vp, i, cp, vp, i]
    L.aclblasCgemm3m.argtypes = gemm
    L.aclblasCgemm.argtypes = gemm
    L.aclblasCgemm3mEx.argtypes = gemm + [ctypes.POINTER(Config)]
    L.aclblasCgemm3mNoAlloc.argtypes = gemm
    i64 = ctypes.c_int64
    L.aclblasCgemm3mStridedBatched.argtypes = [vp, i, i, i, i, i, cp, vp, i, i64, vp, i, i64, cp, vp, i, i64, i]
    L.aclblasCgemm3mGetWorkspaceSize.argtypes = [vp, i, i, i, i, i, cp, cp, ctypes.POINTER(Config),
                                                 ctypes.POINTER(sz)]
    ip = ctypes.POINTER(i)
    L.aclblasCgemm3mGroupedBatched.argtypes = [vp, i, ip, ip, ip, ip, ip, cp, ctypes.POINTER(vp), ip,
                                               ctypes.POINTER(vp), ip, cp, ctypes.POINTER(vp), ip, i]
    L.aclblasCgemm3mGroupedBatchedEx.argtypes = [vp, i, ip, ip, ip, ip, ip, cp, ctypes.POINTER(vp), ip,
                                                 ctypes.POINTER(vp), ip, cp, ctypes.POINTER(vp), ip,
                                                 ctypes.POINTER(Config)]
    L.aclblasCgemm3mGroupedBatchedGetWorkspaceSize.argtypes = [vp, i, ip, ip, ip, ip, ip, i,
                                                               ctypes.POINTER(sz)]
    L.aclblasCgemm3mResidual.argtypes = [vp, i, i, i, i, i, cp, vp, i, vp, i, cp, vp, vp, i,
                                         ctypes.POINTER(ctypes.c_double)]
    L.aclblasCgemm3mGetLastPlan.argtypes = [vp, i, ctypes.POINTER(ctypes.c_longlong)]
    L.aclblasSetLogCallback.argtypes = [LOG_CB]
    L.aclblasCgemm3mSetTiming.argtypes = [vp, i]
    L.aclblasCgemm3mGetTiming.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)]
    for name in EXPORTS:
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def complex_arg(z):
    if z is None:
        return None
    z = complex(*z) if isinstance(z, (tuple, list)) else complex(z)
    return ctypes.byref(Complex(z.real, z.imag))
