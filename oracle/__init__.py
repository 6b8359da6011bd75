"""TEST INFRASTRUCTURE ONLY: the CPU checker for aclblasCgemm3m.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package,
and only to check or time the CPU baseline; the product path never does.

Contents (each pinned by tests/test_oracle.py):
  cgemm()          Netlib reference-BLAS cgemm loop order + the reference golden's
                   k==0/alpha==0 handling (ops-blas/test/gemm/cgemm3m/cgemm3m_golden.h:20-49)
  zgemm()          FP64 complex GEMM of the FP32 inputs (error report)
  verify()         the reference verifier verdict (cgemm3m_verify.h:22-84)
  fill()           counter-based uniform[-1,1) stream, same values as oracle.c
  reltol_check()   |x-g| <= 1e-5*|g| + 1e-5*max|g| per real component
  bruteforce()     exact rational sums for tiny sizes (pins cgemm/zgemm)
"""
import ctypes
import os
import subprocess
from fractions import Fraction

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
OP = {"N": 111, "T": 112, "C": 113}
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        i = ctypes.c_int
        L.oracle_cgemm.argtypes = [i, i, i, i, i, fp, fp, i, fp, i, fp, fp, i, i]
        L.oracle_zgemm.argtypes = [i, i, i, i, i, fp, fp, i, fp, i, fp, fp, i, dp, i, i]
        L.oracle_verify.argtypes = [fp, fp, fp, i, i, i, dp]
        L.oracle_verify_component.argtypes = [fp, fp, ctypes.c_longlong, dp]
        L.oracle_fill.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, fp]
        L.oracle_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_uniform.restype = ctypes.c_float
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _op(x):
    return OP[x] if isinstance(x, str) else int(x)


def _c64(a):
    a = np.ascontiguousarray(a, dtype=np.complex64)
    return a


def cgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, threads=0):
    """Netlib-order cgemm. A, B, C are 1-D complex64 column-major storage; returns new C."""
    al = np.array([np.float32(np.real(alpha)), np.float32(np.imag(alpha))], dtype=np.float32)
    be = np.array([np.float32(np.real(beta)), np.float32(np.imag(beta))], dtype=np.float32)
    A = _c64(A) if A is not None else np.zeros(1, np.complex64)
    B = _c64(B) if B is not None else np.zeros(1, np.complex64)
    out = _c64(C).copy()
    lib().oracle_cgemm(_op(ta), _op(tb), m, n, k, _fp(al), _fp(A.view(np.float32)), lda,
                       _fp(B.view(np.float32)), ldb, _fp(be), _fp(out.view(np.float32)), ldc, threads)
    return out


def zgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, threads=0):
    """FP64 reference: returns complex128 m x n (column-major, ld=m) = alpha*op(A)op(B)+beta*C."""
    al = np.array([np.float32(np.real(alpha)), np.float32(np.imag(alpha))], dtype=np.float32)
    be = np.array([np.float32(np.real(beta)), np.float32(np.imag(beta))], dtype=np.float32)
    A = _c64(A)
    B = _c64(B)
    C = _c64(C) if C is not None else np.zeros(1, np.complex64)
    out = np.zeros(max(m, 1) * max(n, 1), dtype=np.complex128)
    lib().oracle_zgemm(_op(ta), _op(tb), m, n, k, _fp(al), _fp(A.view(np.float32)), lda,
                       _fp(B.view(np.float32)), ldb, _fp(be), _fp(C.view(np.float32)), ldc,
                       _dp(out.view(np.float64)), max(m, 1), threads)
    return out[: m * n]


def verify(actual, golden, initial, m, n, ldc):
    """Reference verifier verdict. Returns (passed, stats dict)."""
    a = _c64(actual)
    g = _c64(golden)
    i0 = _c64(initial)
    st = np.zeros(7, dtype=np.float64)
    ok = lib().oracle_verify(_fp(a.view(np.float32)), _fp(g.view(np.float32)), _fp(i0.view(np.float32)),
                             m, n, ldc, _dp(st))
    keys = ["real_pass", "imag_pass", "padding_ok", "real_max_err", "imag_max_err",
            "real_match_ratio", "imag_match_ratio"]
    return bool(ok), dict(zip(keys, st.tolist()))


def verify_component(actual, golden):
    a = np.ascontiguousarray(actual, dtype=np.float32)
    g = np.ascontiguousarray(golden, dtype=np.float32)
    st = np.zeros(7, dtype=np.float64)
    ok = lib().oracle_verify_component(_fp(a), _fp(g), a.size, _dp(st))
    return bool(ok), st


_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G = np.uint64(0x9E3779B97F4A7C15)
_H = np.uint64(0xD1B54A32D192ED03)


def uniform(seed, start, count):
    """Same stream as oracle.c's oracle_fill: float32 uniform on [-1, 1)."""
    with np.errstate(over="ignore"):
        idx = np.arange(start, start + count, dtype=np.uint64) + np.uint64(1)
        z = np.uint64(seed) * _G + idx * _H
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(40)).astype(np.float32)
    return u * np.float32(2.0 ** -23) - np.float32(1.0)


def fill(rows, cols, ld, seed):
    """Column-major complex64 storage of ld*cols elements (padding rows filled too)."""
    count = max(ld, 1) * max(cols, 0)
    f = uniform(seed, 0, 2 * count)
    return f.view(np.complex64).copy()


def reltol_check(x, g, rtol=1e-5, atol_frac=1e-5):
    """Per real component: |x-g| <= rtol*|g| + atol_frac*max|g|. Returns (ok, worst ratio)."""
    xv = np.asarray(x, dtype=np.complex64).view(np.float32).astype(np.float64)
    gv = np.asarray(g, dtype=np.complex64).view(np.float32).astype(np.float64)
    if gv.size == 0:
        return True, 0.0
    fin = np.isfinite(gv) & np.isfinite(xv)
    if not np.array_equal(np.isfinite(gv), np.isfinite(xv)):
        return False, float("inf")
    scale = np.max(np.abs(gv[fin])) if fin.any() else 0.0
    bound = rtol * np.abs(gv[fin]) + atol_frac * scale
    err = np.abs(xv[fin] - gv[fin])
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.where(bound > 0, err / np.where(bound > 0, bound, 1), np.where(err > 0, np.inf, 0.0))
    worst = float(ratio.max()) if ratio.size else 0.0
    return worst <= 1.0, worst


def op_matrix(X, rows, cols, ld, op):
    """Dense op(X) as complex128 from column-major storage of a rows x cols matrix."""
    M = np.asarray(X, dtype=np.complex64)[: ld * cols].reshape(cols, ld).T[:rows, :].astype(np.complex128)
    if op in ("T", 112):
        M = M.T
    elif op in ("C", 113):
        M = M.T.conj()
    return M


def errors_vs(x, g):
    """max/mean elementwise |x-g|/|g| (g != 0), normwise max|x-g|/max|g| (complex modulus)."""
    x = np.asarray(x).astype(np.complex128).ravel()
    g = np.asarray(g).astype(np.complex128).ravel()
    d = np.abs(x - g)
    ag = np.abs(g)
    nz = ag > 0
    el = d[nz] / ag[nz]
    return {
        "max_rel": float(el.max()) if el.size else 0.0,
        "mean_rel": float(el.mean()) if el.size else 0.0,
        "normwise": float(d.max() / ag.max()) if ag.size and ag.max() > 0 else 0.0,
    }


def bruteforce(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    """Exact rational result (as (re, im) Fraction pairs, column-major m x n) for tiny sizes."""
    def F(v):
        return Fraction(float(np.float32(v)))

    def get(X, ld, r, c):
        v = X[c * ld + r]
        return F(v.real), F(v.imag)

    ar, ai = F(np.real(alpha)), F(np.imag(alpha))
    br, bi = F(np.real(beta)), F(np.imag(beta))
    out = []
    for j in range(n):
        for i in range(m):
            sr = si = Fraction(0)
            for l in range(k):
                xr, xi = get(A, lda, i, l) if ta in ("N", 111) else get(A, lda, l, i)
                if ta in ("C", 113):
                    xi = -xi
                yr, yi = get(B, ldb, l, j) if tb in ("N", 111) else get(B, ldb, j, l)
                if tb in ("C", 113):
                    yi = -yi
                sr += xr * yr - xi * yi
                si += xr * yi + xi * yr
            vr = ar * sr - ai * si
            vi = ar * si + ai * sr
            if not (br == 0 and bi == 0):
                cr, ci = get(C, ldc, i, j)
                vr += br * cr - bi * ci
                vi += br * ci + bi * cr
            out.append((vr, vi))
    return out
