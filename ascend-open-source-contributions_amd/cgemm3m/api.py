"""Host-side mirror of the reference's aclblasCgemm3m calling convention.

``Handle`` wraps aclblasCreate/Destroy/SetStream (ops-blas/include/cann_ops_blas.h:32-110);
``cgemm3m`` forwards one call with the reference's argument meaning and status codes
(cann_ops_blas.h:497-500). Matrices are device pointers (ints) to column-major complex64
storage, exactly as the C ABI takes them; ``torch_ptr`` turns a torch tensor into one.
"""
import ctypes

from . import _lib


class Cgemm3mError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__("%s returned %s (%d)" % (what, _lib.STATUS.get(status, "?"), status))


def _op(t):
    return _lib.OP[t] if isinstance(t, str) else int(t)


def torch_ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class Handle:
    """One aclblas handle (one host thread at a time, like the reference)."""

    def __init__(self, stream=None):
        self.L = _lib.load()
        h = ctypes.c_void_p()
        st = self.L.aclblasCreate(ctypes.byref(h))
        if st != _lib.SUCCESS:
            raise Cgemm3mError(st, "aclblasCreate")
        self.h = h
        if stream is not None:
            self.set_stream(stream)

    def set_stream(self, stream):
        """stream: a torch.cuda.Stream, a raw hipStream_t int, or None (default stream)."""
        raw = getattr(stream, "cuda_stream", stream)
        st = self.L.aclblasSetStream(self.h, ctypes.c_void_p(raw or 0))
        if st != _lib.SUCCESS:
            raise Cgemm3mError(st, "aclblasSetStream")

    def close(self):
        if self.h is not None and self.h.value:
            self.L.aclblasDestroy(self.h)
        self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, enable):
        st = self.L.aclblasCgemm3mSetTiming(self.h, 1 if enable else 0)
        if st != _lib.SUCCESS:
            raise Cgemm3mError(st, "aclblasCgemm3mSetTiming")

    def timing(self):
        """{stage: summed ms} over the calls recorded since the last read, plus 'calls'."""
        ms = (ctypes.c_double * 6)()
        calls = ctypes.c_longlong()
        st = self.L.aclblasCgemm3mGetTiming(self.h, ms, ctypes.byref(calls))
        if st != _lib.SUCCESS:
            raise Cgemm3mError(st, "aclblasCgemm3mGetTiming")
        out = dict(zip(_lib.STAGES, list(ms)))
        out["calls"] = calls.value
        return out

    def last_plan(self, sync=False):
        info = (ctypes.c_longlong * 8)()
        st = self.L.aclblasCgemm3mGetLastPlan(self.h, 1 if sync else 0, info)
        if st != _lib.SUCCESS:
            raise Cgemm3mError(st, "aclblasCgemm3mGetLastPlan")
        keys = ["path", "tile", "parts", "correction", "precision", "identity", "flagged", "workspace"]
        return dict(zip(keys, list(info)))


def config(precision="fp32", tile=0, k_split=0, correction=0, identity=0, variant=0):
    prec = _lib.PRECISION[precision] if isinstance(precision, str) else int(precision)
    return _lib.Config(prec, tile, k_split, correction, identity, variant)


def cgemm3m(handle, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, cfg=None, check=True):
    """aclblasCgemm3m (or aclblasCgemm3mEx when cfg is given). Returns the status code."""
    L = handle.L
    args = [handle.h if handle is not None else None, _op(transa), _op(transb), m, n, k,
            _lib.complex_arg(alpha), A, lda, B, ldb, _lib.complex_arg(beta), C, ldc]
    if cfg is None:
        st = L.aclblasCgemm3m(*args)
        what = "aclblasCgemm3m"
    else:
        st = L.aclblasCgemm3mEx(*(args + [ctypes.byref(cfg)]))
        what = "aclblasCgemm3mEx"
    if check and st != _lib.SUCCESS:
        raise Cgemm3mError(st, what)
    return st


def cgemm(handle, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, check=True):
    """aclblasCgemm: the four-product complex GEMM. Returns the status code."""
    st = handle.L.aclblasCgemm(handle.h if handle is not None else None, _op(transa), _op(transb), m, n, k,
                               _lib.complex_arg(alpha), A, lda, B, ldb, _lib.complex_arg(beta), C, ldc)
    if check and st != _lib.SUCCESS:
        raise Cgemm3mError(st, "aclblasCgemm")
    return st


def strided_batched(handle, transa, transb, m, n, k, alpha, A, lda, stride_a, B, ldb, stride_b, beta, C, ldc,
                    stride_c, batch, check=True):
    """aclblasCgemm3mStridedBatched (strides in complex elements). Returns the status code."""
    L = handle.L
    st = L.aclblasCgemm3mStridedBatched(handle.h if handle is not None else None, _op(transa), _op(transb), m, n,
                                        k, _lib.complex_arg(alpha), A, lda, stride_a, B, ldb, stride_b,
                                        _lib.complex_arg(beta), C, ldc, stride_c, batch)
    if check and st != _lib.SUCCESS:
        raise Cgemm3mError(st, "aclblasCgemm3mStridedBatched")
    return st


def residual(handle, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, Cinit, C, ldc):
    """FP64 error report of C against alpha*op(A)op(B)+beta*Cinit (synchronises)."""
    out = (ctypes.c_double * 4)()
    st = handle.L.aclblasCgemm3mResidual(handle.h, _op(transa), _op(transb), m, n, k,
                                         _lib.complex_arg(alpha), A, lda, B, ldb, _lib.complex_arg(beta),
                                         Cinit, C, ldc, out)
    if st != _lib.SUCCESS:
        raise Cgemm3mError(st, "aclblasCgemm3mResidual")
    return {"max_rel": out[0], "mean_rel": out[1], "normwise": out[2], "max_ref": out[3]}


def workspace_size(handle, transa, transb, m, n, k, alpha, beta, cfg=None):
    n_bytes = ctypes.c_size_t()
    st = handle.L.aclblasCgemm3mGetWorkspaceSize(handle.h, _op(transa), _op(transb), m, n, k,
                                                 _lib.complex_arg(alpha), _lib.complex_arg(beta),
                                                 ctypes.byref(cfg) if cfg is not None else None,
                                                 ctypes.byref(n_bytes))
    if st != _lib.SUCCESS:
        raise Cgemm3mError(st, "aclblasCgemm3mGetWorkspaceSize")
    return n_bytes.value


def _arrays(probs):
    cnt = len(probs)
    I = ctypes.c_int * cnt
    CX = _lib.Complex * cnt
    VP = ctypes.c_void_p * cnt

    def ints(key, f=int):
        return I(*[f(p[key]) for p in probs])

    def cx(key):
        return CX(*[_lib.Complex(complex(p[key]).real, complex(p[key]).imag) for p in probs])

    def ptrs(key):
        return VP(*[(p[key].value if isinstance(p[key], ctypes.c_void_p) else p[key]) for p in probs])

    return dict(ta=ints("ta", _op), tb=ints("tb", _op), m=ints("m"), n=ints("n"), k=ints("k"),
                alpha=cx("alpha"), beta=cx("beta"), A=ptrs("A"), lda=ints("lda"), B=ptrs("B"), ldb=ints("ldb"),
                C=ptrs("C"), ldc=ints("ldc"))


def grouped(handle, probs, precision="fp32", check=True, cfg=None):
    """aclblasCgemm3mGroupedBatched (aclblasCgemm3mGroupedBatchedEx when cfg is given) over a
    list of problem dicts with keys ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc
    (A/B/C device pointers)."""
    a = _arrays(probs)
    args = [handle.h, len(probs), a["ta"], a["tb"], a["m"], a["n"], a["k"], a["alpha"], a["A"], a["lda"], a["B"],
            a["ldb"], a["beta"], a["C"], a["ldc"]]
    if cfg is None:
        prec = _lib.PRECISION[precision] if isinstance(precision, str) else int(precision)
        st = handle.L.aclblasCgemm3mGroupedBatched(*(args + [prec]))
        what = "aclblasCgemm3mGroupedBatched"
    else:
        st = handle.L.aclblasCgemm3mGroupedBatchedEx(*(args + [ctypes.byref(cfg)]))
        what = "aclblasCgemm3mGroupedBatchedEx"
    if check and st != _lib.SUCCESS:
        raise Cgemm3mError(st, what)
    return st


def grouped_workspace_size(handle, probs, precision="fp32"):
    a = _arrays([dict(p, A=None, B=None, C=None, alpha=1, beta=0, lda=1, ldb=1, ldc=1) for p in probs])
    prec = _lib.PRECISION[precision] if isinstance(precision, str) else int(precision)
    n_bytes = ctypes.c_size_t()
    st = handle.L.aclblasCgemm3mGroupedBatchedGetWorkspaceSize(handle.h, len(probs), a["ta"], a["tb"], a["m"],
                                                               a["n"], a["k"], prec, ctypes.byref(n_bytes))
    if st != _lib.SUCCESS:
        raise Cgemm3mError(st, "aclblasCgemm3mGroupedBatchedGetWorkspaceSize")
    return n_bytes.value
