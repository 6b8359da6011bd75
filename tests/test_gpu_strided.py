"""GPU tests of aclblasCgemm3mStridedBatched against the Netlib-order oracle, batch by batch.

Contract (ops-blas/blas/gemm/README.md, section aclblasCgemm3mStridedBatched): batch i uses
A + i*strideA, B + i*strideB, C + i*strideC; zero and negative strides are allowed; alpha == 0
only scales by beta; overlapping output batches are updated in batch order; reruns are
byte-identical. Every case checks each batch at the 1e-5 mixed tolerance plus the reference
verifier, every element outside the batches' outputs untouched, and a byte-identical rerun.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from helpers import to_dev

pytestmark = pytest.mark.gpu

ALPHA, BETA = complex(0.7, -0.3), complex(0.2, 0.9)


def _ptr(t, offset):
    return ctypes.c_void_p(t.data_ptr() + 8 * offset)


def _expected(ta, tb, m, n, k, alpha, A, lda, a0, sa, B, ldb, b0, sb, beta, C, ldc, c0, sc, batch):
    """Apply the oracle batch after batch to one flat copy of C (batch order matters only
    when outputs overlap, and is what the contract asks for then)."""
    out = C.copy()
    ext_a = (k if ta == "N" else m) * lda
    ext_b = (n if tb == "N" else k) * ldb
    ext_c = (n - 1) * ldc + m
    for i in range(batch):
        ao, bo, co = a0 + i * sa, b0 + i * sb, c0 + i * sc
        Ai = A[ao:ao + ext_a] if A is not None else None
        Bi = B[bo:bo + ext_b] if B is not None else None
        g = oracle.cgemm(ta, tb, m, n, k, alpha, Ai, lda, Bi, ldb, beta, out[co:co + ext_c], ldc)
        out[co:co + ext_c] = g[:ext_c]
    return out


def _run(h, ta, tb, m, n, k, alpha, A, lda, a0, sa, B, ldb, b0, sb, beta, C, ldc, c0, sc, batch, label):
    import torch

    import cgemm3m

    print("case %s %s%s m=%d n=%d k=%d strides=(%d,%d,%d) batch=%d" % (label, ta, tb, m, n, k, sa, sb, sc, batch))
    sys.stdout.flush()
    dA = to_dev(A) if A is not None else None
    dB = to_dev(B) if B is not None else None
    outs = []
    for _ in range(2):
        dC = to_dev(C)
        cgemm3m.strided_batched(h, ta, tb, m, n, k, alpha, _ptr(dA, a0) if dA is not None else None, lda, sa,
                                _ptr(dB, b0) if dB is not None else None, ldb, sb, beta, _ptr(dC, c0), ldc, sc,
                                batch)
        torch.cuda.synchronize()
        outs.append(dC.cpu().numpy())
    same = outs[0].tobytes() == outs[1].tobytes()
    assert same, "%s: rerun differs" % label
    return outs[0]


def _check(out, gold, C, m, n, ldc, c0, sc, batch, label):
    """Each batch's m x n block at reltol + verifier; everything else byte-identical to C."""
    touched = np.zeros(C.shape[0], dtype=bool)
    for i in range(batch):
        co = c0 + i * sc
        for j in range(n):
            touched[co + j * ldc: co + j * ldc + m] = True
    for i in range(batch):
        co = c0 + i * sc
        ext = (n - 1) * ldc + m
        blk, ref = out[co:co + ext], gold[co:co + ext]
        r_out = np.stack([blk[j * ldc: j * ldc + m] for j in range(n)], axis=1)
        r_ref = np.stack([ref[j * ldc: j * ldc + m] for j in range(n)], axis=1)
        ok, worst = oracle.reltol_check(r_out, r_ref)
        assert ok, "%s batch %d: reltol worst %.3g" % (label, i, worst)
        verdict, stats = oracle.verify(r_out.T.reshape(-1), r_ref.T.reshape(-1), r_ref.T.reshape(-1), m, n, m)
        assert verdict, "%s batch %d: verifier %s" % (label, i, stats)
    untouched = out[~touched].tobytes() == C[~touched].tobytes()
    assert untouched, "%s: elements outside the outputs were written" % label


def _fill(count, seed):
    return oracle.fill(count, 1, count, seed)


def test_strided_disjoint_padded(gpu):
    """Disjoint batches with gaps between them, ld padding, (C, T)."""
    ta, tb, m, n, k, batch = "C", "T", 150, 90, 200, 5
    lda, ldb, ldc = k + 3, n + 2, m + 5
    sa, sb, sc = lda * m + 11, ldb * k + 7, ldc * n + 13
    A, B, C = _fill(sa * batch, 20270001), _fill(sb * batch, 20270002), _fill(sc * batch, 20270003)
    out = _run(gpu, ta, tb, m, n, k, ALPHA, A, lda, 0, sa, B, ldb, 0, sb, BETA, C, ldc, 0, sc, batch, "disjoint")
    gold = _expected(ta, tb, m, n, k, ALPHA, A, lda, 0, sa, B, ldb, 0, sb, BETA, C, ldc, 0, sc, batch)
    _check(out, gold, C, m, n, ldc, 0, sc, batch, "disjoint")


def test_strided_zero_and_negative_input_strides(gpu):
    """strideA = 0 (one A for all batches), strideB < 0 (B walks backwards), (N, C)."""
    ta, tb, m, n, k, batch = "N", "C", 260, 300, 333, 4
    lda, ldb, ldc = m, n, m
    sb = -(ldb * k)
    sc = ldc * n
    A = _fill(lda * k, 20270011)
    B = _fill(-sb * batch, 20270012)
    C = _fill(sc * batch, 20270013)
    b0 = -sb * (batch - 1)  # batch 0 uses the last block, batch i the block before it
    out = _run(gpu, ta, tb, m, n, k, ALPHA, A, lda, 0, 0, B, ldb, b0, sb, BETA, C, ldc, 0, sc, batch, "zero-neg")
    gold = _expected(ta, tb, m, n, k, ALPHA, A, lda, 0, 0, B, ldb, b0, sb, BETA, C, ldc, 0, sc, batch)
    _check(out, gold, C, m, n, ldc, 0, sc, batch, "zero-neg")


@pytest.mark.parametrize("sc_kind", ["zero", "half"])
def test_strided_overlapping_outputs_in_batch_order(gpu, sc_kind):
    """strideC = 0 (every batch updates the same C) and a stride of half a column block:
    outputs overlap, so batch i must see batch i-1's result (beta != 0 makes order visible)."""
    ta, tb, m, n, k, batch = "T", "N", 96, 64, 130, 3
    lda, ldb, ldc = k, k, m
    sa, sb = lda * m, ldb * n
    sc = 0 if sc_kind == "zero" else (n // 2) * ldc
    A, B = _fill(sa * batch, 20270021), _fill(sb * batch, 20270022)
    total_c = (n - 1) * ldc + m + sc * (batch - 1)
    C = _fill(total_c, 20270023)
    out = _run(gpu, ta, tb, m, n, k, ALPHA, A, lda, 0, sa, B, ldb, 0, sb, BETA, C, ldc, 0, sc, batch, "overlap")
    gold = _expected(ta, tb, m, n, k, ALPHA, A, lda, 0, sa, B, ldb, 0, sb, BETA, C, ldc, 0, sc, batch)
    ok, worst = oracle.reltol_check(out, gold)
    assert ok, "overlap: reltol worst %.3g" % worst


def test_strided_alpha_zero_scales_by_beta(gpu):
    """alpha == 0: A and B may be null and are not read; C_i = beta*C_i per batch, and with
    strideC == 0 the single C is scaled once per batch."""
    m, n, k, batch = 70, 40, 50, 3
    ldc = m + 1
    sc = ldc * n
    C = _fill(sc * batch, 20270031)
    out = _run(gpu, "N", "N", m, n, k, 0, None, m, 0, 0, None, k, 0, 0, BETA, C, ldc, 0, sc, batch, "alpha0")
    gold = _expected("N", "N", m, n, k, 0, None, m, 0, 0, None, k, 0, 0, BETA, C, ldc, 0, sc, batch)
    _check(out, gold, C, m, n, ldc, 0, sc, batch, "alpha0")
    out0 = _run(gpu, "N", "N", m, n, k, 0, None, m, 0, 0, None, k, 0, 0, BETA, C, ldc, 0, 0, batch, "alpha0-sc0")
    gold0 = _expected("N", "N", m, n, k, 0, None, m, 0, 0, None, k, 0, 0, BETA, C, ldc, 0, 0, batch)
    ok, worst = oracle.reltol_check(out0, gold0)
    assert ok, worst


def test_strided_matches_per_batch_calls(gpu):
    """A uniform batch of 16 at 512^3 (one grouped launch) equals 16 aclblasCgemm3mEx calls with
    the member plan, byte for byte, and the oracle on two sampled batches. The member plan
    of a 512^3 grouped member is the untuned default (128 tiles, no K split, no correction):
    the plan table holds no grouped entry for k in [512, 1024)."""
    import torch

    import cgemm3m

    m = n = k = 512
    batch = 16
    s = m * k
    A, B, C = _fill(s * batch, 20270041), _fill(s * batch, 20270042), _fill(s * batch, 20270043)
    out = _run(gpu, "N", "N", m, n, k, ALPHA, A, m, 0, s, B, k, 0, s, BETA, C, m, 0, s, batch, "uniform16")
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    cfg = cgemm3m.config(tile=128, k_split=1, correction=-1, variant=3)
    for i in range(batch):
        cgemm3m.cgemm3m(gpu, "N", "N", m, n, k, ALPHA, _ptr(dA, i * s), m, _ptr(dB, i * s), k, BETA,
                        _ptr(dC, i * s), m, cfg)
    torch.cuda.synchronize()
    loop = dC.cpu().numpy()
    same = loop.tobytes() == out.tobytes()
    assert same, "strided batch differs from per-batch calls with the member plan"
    for i in (0, batch - 1):
        g = oracle.cgemm("N", "N", m, n, k, ALPHA, A[i * s:(i + 1) * s], m, B[i * s:(i + 1) * s], k, BETA,
                         C[i * s:(i + 1) * s], m)
        ok, worst = oracle.reltol_check(out[i * s:(i + 1) * s], g)
        assert ok, worst
