"""Shared test plumbing: seeded inputs, one GPU call through the C ABI, and the checks.

Inputs come from the oracle's counter-based generator (A, B, C use seeds s, s+1, s+2, as
ops-blas/test/gemm/cgemm3m/arch35/cgemm3m_test.cpp:276-280 does with its own fills).
"""
import sys

import numpy as np

import oracle


def stored_dims(t, rows, cols):
    """(rows, cols) of the stored matrix whose op() is rows x cols."""
    return (rows, cols) if t == "N" else (cols, rows)


def inputs(ta, tb, m, n, k, lda, ldb, ldc, seed):
    ar, ac = stored_dims(ta, m, k)
    br, bc = stored_dims(tb, k, n)
    A = oracle.fill(ar, ac, lda, seed)
    B = oracle.fill(br, bc, ldb, seed + 1)
    C = oracle.fill(m, n, ldc, seed + 2)
    return A, B, C


def to_dev(x):
    import torch

    if x is None or x.size == 0:
        return None
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def ptr(t):
    import ctypes

    return None if t is None else ctypes.c_void_p(t.data_ptr())


def run_gpu(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, cfg=None, dA=None, dB=None):
    """One aclblasCgemm3m(Ex) call on device copies; returns (status, C after the call)."""
    import torch

    import cgemm3m

    dA = dA if dA is not None else to_dev(A)
    dB = dB if dB is not None else to_dev(B)
    dC = to_dev(C)
    st = cgemm3m.cgemm3m(h, ta, tb, m, n, k, alpha, ptr(dA), lda, ptr(dB), ldb, beta, ptr(dC), ldc, cfg,
                         check=False)
    torch.cuda.synchronize()
    out = dC.cpu().numpy() if dC is not None else np.zeros(0, np.complex64)
    return st, out


def region(x, m, n, ld):
    """The m x n logical part of column-major storage."""
    return np.asarray(x)[: ld * n].reshape(n, ld)[:, :m].T.copy()


def check_against_oracle(out, gold, initial, m, n, ldc, label=""):
    """reltol 1e-5 (+1e-5*max) per component, the reference verifier, padding untouched."""
    ok, worst = oracle.reltol_check(region(out, m, n, ldc), region(gold, m, n, ldc))
    verdict, stats = oracle.verify(out, gold, initial, m, n, ldc)
    print("%s reltol_worst=%.3g verifier=%s %s" % (label, worst, verdict, stats))
    sys.stdout.flush()
    assert ok, "%s: reltol worst ratio %.3g" % (label, worst)
    assert verdict, "%s: reference verifier failed %s" % (label, stats)
    pad = np.asarray(out)[: ldc * n].reshape(n, ldc)[:, m:]
    pad0 = np.asarray(initial)[: ldc * n].reshape(n, ldc)[:, m:]
    same = pad.tobytes() == pad0.tobytes()
    assert same, "%s: padding rows were written" % label
