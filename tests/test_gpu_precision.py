"""GPU tests of the split-precision forms: bf16x3 (hi+lo BF16, 3 MFMA products per real
product), bf16x6 (three BF16 pieces, 6 products) and FP16 with a power-of-two scale per
128 x 128 block. None has a reference counterpart, so each is pinned against the FP64
complex GEMM of the same inputs (the GPU residual kernel, itself checked against the CPU
FP64 oracle in test_gpu_parity.py) at a bound set by its arithmetic, plus the Netlib
oracle at a loosened reltol, plus byte-identical reruns.
"""
import sys

import numpy as np
import pytest

import oracle
from helpers import inputs, ptr, run_gpu, to_dev

pytestmark = pytest.mark.gpu

# normwise bound max|c-g|/max|g| against FP64 for k ~ 1e3 uniform[-1,1] inputs
BOUND = {"fp32": 2e-6, "bf16x3": 6e-5, "bf16x6": 2e-6, "fp16": 4e-3}
ALPHA, BETA = complex(0.8, 0.6), complex(0.5, 0.0)


def _run(h, prec, ta, tb, m, n, k, seed, tile=0, parts=0, scale_rows=False):
    import cgemm3m

    print("case %s %s%s m=%d n=%d k=%d tile=%d parts=%d scaled=%s" % (prec, ta, tb, m, n, k, tile, parts,
                                                                       scale_rows))
    sys.stdout.flush()
    lda = (m if ta == "N" else k)
    ldb = (k if tb == "N" else n)
    A, B, C = inputs(ta, tb, m, n, k, lda, ldb, m, seed)
    if scale_rows:
        # rows of op(A) scaled by 10^u, u ~ U[-3, 3]
        u = oracle.uniform(seed + 99, 0, m).astype(np.float64) * 3.0
        s = (10.0 ** u).astype(np.float32)
        Ad = A.reshape(-1, lda).copy()  # stored columns
        if ta == "N":
            Ad[:, :m] *= s[None, :]
        else:
            Ad[:m, :] *= s[:, None]
        A = Ad.reshape(-1).astype(np.complex64)
    cfg = cgemm3m.config(prec, tile=tile, k_split=parts)
    dA, dB = to_dev(A), to_dev(B)
    st, out = run_gpu(h, ta, tb, m, n, k, ALPHA, A, lda, B, ldb, BETA, C, m, cfg, dA, dB)
    assert st == 0
    plan = h.last_plan(sync=True)
    st2, out2 = run_gpu(h, ta, tb, m, n, k, ALPHA, A, lda, B, ldb, BETA, C, m, cfg, dA, dB)
    same = st2 == 0 and out2.tobytes() == out.tobytes()
    assert same, "rerun differs"
    dC0, dOut = to_dev(C), to_dev(out)
    res = cgemm3m.residual(h, ta, tb, m, n, k, ALPHA, ptr(dA), lda, ptr(dB), ldb, BETA, ptr(dC0), ptr(dOut), m)
    print("plan", plan, "err", res)
    return out, res, (A, B, C, lda, ldb)


@pytest.mark.parametrize("prec", ["bf16x3", "bf16x6", "fp16"])
@pytest.mark.parametrize("ops", ["NN", "CT", "TC"])
def test_precision_vs_fp64(gpu, prec, ops):
    out, res, (A, B, C, lda, ldb) = _run(gpu, prec, ops[0], ops[1], 384, 320, 1024, 20261100 + len(prec))
    assert res["normwise"] <= BOUND[prec], res
    if prec == "bf16x6":
        gold = oracle.cgemm(ops[0], ops[1], 384, 320, 1024, ALPHA, A, lda, B, ldb, BETA, C, 384)
        ok, worst = oracle.reltol_check(out, gold)
        assert ok, worst


@pytest.mark.parametrize("prec", ["bf16x3", "bf16x6", "fp16"])
def test_precision_ksplit_and_tiles(gpu, prec):
    for tile, parts in ((64, 1), (128, 1), (64, 3), (128, 2)):
        _, res, _ = _run(gpu, prec, "N", "C", 200, 260, 777, 20261200, tile=tile, parts=parts)
        assert res["normwise"] <= BOUND[prec], res


def test_fp16_block_exponents_mixed_and_zero(gpu):
    """FP16 block exponents that differ along k: columns 128..255 of A scaled by 2^20 in rows
    0..255 only (so A's first two 128-row blocks have mixed exponents and the rest uniform ones),
    columns 384..511 zero (an all-zero block, whose exponent is ignored) and columns 640..767
    scaled by 3 (an exponent one higher, snapped to the row block's). The separate-products
    form then splits its workgroups between the uniform launch and the rescaling launch; every FP16
    form without a K split (fused 4 and 8 waves, separate products, its LDS form) gives the same
    bytes, and (also with a K split) each row band (the mixed rows 0..255, the uniform rest) stays within the
    FP16 bound against FP64 on its own: a rescale applied wrongly would put the unscaled
    contributions of rows 0..255 off by 2^20."""
    import cgemm3m

    m, n, k = 700, 600, 1000
    print("case fp16 mixed exponents m=%d n=%d k=%d" % (m, n, k))
    sys.stdout.flush()
    A, B, C = inputs("N", "N", m, n, k, m, k, m, 20261500)
    cols = A.reshape(k, m).copy()  # column-major: row l of this view is column l of A
    cols[128:256, :256] *= np.float32(2.0 ** 20)
    cols[384:512, :] = 0
    cols[640:768, :] *= np.float32(3.0)  # one octave up: snapped to the row block's exponent
    A = cols.reshape(-1)
    dA, dB = to_dev(A), to_dev(B)
    outs = {}
    for variant, parts in ((1, 1), (2, 1), (4, 1), (6, 1), (1, 3)):
        cfg = cgemm3m.config("fp16", tile=128, k_split=parts, variant=variant, correction=-1)
        st, out = run_gpu(gpu, "N", "N", m, n, k, ALPHA, A, m, B, k, BETA, C, m, cfg, dA, dB)
        assert st == 0
        plan = gpu.last_plan(sync=True)
        assert plan["path"] == (4 if variant in (4, 6) else (2 if parts > 1 else 1)), plan
        outs[(variant, parts)] = out
    first = outs[(1, 1)]
    differ = {key: int(np.count_nonzero(out.view(np.uint64) != first.view(np.uint64)))
              for key, out in outs.items() if key[1] == 1}  # a K split sums in another order
    print("outputs differing from the fused kernel's, per form:", differ)
    assert not any(differ.values()), differ
    dC0 = to_dev(C)
    for key in ((1, 1), (1, 3)):
        dOut = to_dev(outs[key])
        for r0, rows in ((0, 256), (256, m - 256)):
            res = cgemm3m.residual(gpu, "N", "N", rows, n, k, ALPHA, ptr(dA[r0:]), m, ptr(dB), k, BETA,
                                   ptr(dC0[r0:]), ptr(dOut[r0:]), m)
            print("form", key, "rows", r0, rows, "err", res)
            assert res["normwise"] <= BOUND["fp16"], res


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "fp16"])
def test_row_scaled_range(gpu, prec):
    """Rows of op(A) spread over 6 decades: FP16 relies on its per-block scales here."""
    _, res, _ = _run(gpu, prec, "N", "N", 256, 256, 512, 20261300, scale_rows=True)
    assert res["normwise"] <= BOUND[prec] * 4, res


def test_bf16x3_error_ordering(gpu):
    """bf16x6 is at least as accurate as FP32-3M and bf16x3 is within its 2^-16 budget."""
    errs = {}
    for prec in ("fp32", "bf16x3", "bf16x6"):
        _, res, _ = _run(gpu, prec, "N", "N", 512, 512, 2048, 20261400)
        errs[prec] = res
    print(errs)
    assert errs["bf16x6"]["normwise"] <= 1.5 * errs["fp32"]["normwise"]
    assert errs["bf16x3"]["normwise"] <= 30 * errs["fp32"]["normwise"]
