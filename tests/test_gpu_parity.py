"""GPU parity: the HIP path through the C ABI against the Netlib-order oracle.

Every case checks (a) |x-g| <= 1e-5|g| + 1e-5 max|g| per real component, (b) the
reference verifier's verdict (cgemm3m_verify.h:22-84) including untouched ldc padding,
and (c) byte-identical output on a second run (cgemm3m_test.cpp:304-307).
"""
import itertools
import sys
import zlib

import numpy as np
import pytest

import oracle
from helpers import check_against_oracle, inputs, ptr, region, run_gpu, to_dev

pytestmark = pytest.mark.gpu

ALPHA2, BETA2 = complex(0.7, -0.3), complex(0.2, 0.9)


def _case(h, ta, tb, m, n, k, alpha, beta, lda, ldb, ldc, seed, cfg=None, label="", twice=True):
    print("case %s %s%s m=%d n=%d k=%d ld=(%d,%d,%d) cfg=%s" % (label, ta, tb, m, n, k, lda, ldb, ldc,
                                                               None if cfg is None else list(cfg._fields_)))
    sys.stdout.flush()
    A, B, C = inputs(ta, tb, m, n, k, lda, ldb, ldc, seed)
    dA, dB = to_dev(A), to_dev(B)
    st, out = run_gpu(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, cfg, dA, dB)
    assert st == 0
    plan = h.last_plan(sync=True)
    print("plan", plan)
    gold = oracle.cgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    check_against_oracle(out, gold, C, m, n, ldc, label)
    if twice:
        st2, out2 = run_gpu(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, cfg, dA, dB)
        same = st2 == 0 and out2.tobytes() == out.tobytes()
        assert same, "second run differs"
    return out, plan


def test_c1_256_nn(gpu):
    """The reference's own headline check: 256^3 NN vs cblas_cgemm, double run byte-identical."""
    _case(gpu, "N", "N", 256, 256, 256, complex(1.0, 0.5), complex(0.5, -0.25), 256, 256, 256, 20260001,
          label="C1")


@pytest.mark.parametrize("ta,tb", list(itertools.product("NTC", "NTC")))
def test_op_sweep_padded(gpu, ta, tb):
    m, n, k = 1000, 1500, 2048
    lda = (m if ta == "N" else k) + 7
    ldb = (k if tb == "N" else n) + 7
    _case(gpu, ta, tb, m, n, k, ALPHA2, BETA2, lda, ldb, m + 7, 20260100 + 3 * "NTC".index(ta) + "NTC".index(tb),
          label="sweep")


def test_square_2048_fused_with_correction(gpu):
    """2048^3 on the fused tile (pinned: no K split) with the correction screen on (auto at
    k >= 1024)."""
    import cgemm3m

    cfg = cgemm3m.config(tile=128, k_split=1)
    out, plan = _case(gpu, "N", "N", 2048, 2048, 2048, ALPHA2, BETA2, 2048, 2048, 2048, 20260200, cfg=cfg,
                      label="sq2048")
    assert plan["path"] == 1 and plan["correction"] == 1


def test_square_4096_sampled_columns(gpu):
    """4096^3: full FP64 residual on the GPU plus the Netlib oracle on 24 sampled columns
    (Netlib columns are independent, so a column subset is the same arithmetic)."""
    import cgemm3m

    m = n = k = 4096
    A, B, C = inputs("N", "N", m, n, k, m, k, m, 20260300)
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    dC0 = to_dev(C)
    cgemm3m.cgemm3m(gpu, "N", "N", m, n, k, ALPHA2, ptr(dA), m, ptr(dB), k, BETA2, ptr(dC), m)
    out = dC.cpu().numpy()
    res = cgemm3m.residual(gpu, "N", "N", m, n, k, ALPHA2, ptr(dA), m, ptr(dB), k, BETA2, ptr(dC0), ptr(dC), m)
    print("4096 residual", res)
    assert res["normwise"] < 1e-5
    cols = np.linspace(0, n - 1, 24).astype(int)
    Bs = np.concatenate([B[j * k:(j + 1) * k] for j in cols])
    Cs = np.concatenate([C[j * m:(j + 1) * m] for j in cols])
    gold = oracle.cgemm("N", "N", m, len(cols), k, ALPHA2, A, m, Bs, k, BETA2, Cs, m)
    outs = np.concatenate([out[j * m:(j + 1) * m] for j in cols])
    ok, worst = oracle.reltol_check(outs, gold)
    assert ok, worst


_TS = {}


def _tall_skinny_inputs(h):
    """256 x 256 x 65536 inputs, the Netlib-order result and Netlib's own FP64 error."""
    import cgemm3m

    if not _TS:
        m = n = 256
        k = 65536
        A, B, C = inputs("N", "N", m, n, k, m, k, m, 20260400)
        gold = oracle.cgemm("N", "N", m, n, k, ALPHA2, A, m, B, k, BETA2, C, m)
        dA, dB, dC0, dG = to_dev(A), to_dev(B), to_dev(C), to_dev(gold)
        err = cgemm3m.residual(h, "N", "N", m, n, k, ALPHA2, ptr(dA), m, ptr(dB), k, BETA2, ptr(dC0), ptr(dG), m)
        _TS.update(A=A, B=B, C=C, gold=gold, dA=dA, dB=dB, dC0=dC0, netlib_err=err)
    return _TS


@pytest.mark.parametrize("parts", [0, 1, 2, 3, 4])
def test_tall_skinny_ksplit(gpu, parts):
    """K-split 1-4 (and the planner's own choice) at k = 65536.

    Netlib's sequential FP32 sum over 65536 terms itself only just meets the reference
    verifier against exact FP64 (about 99.4% of components within (1+|g|)/8192), so two
    independent FP32 chains of this length cannot both meet it against each other. The
    check is therefore: reltol 1e-5 against Netlib (the stated tolerance), FP64 error no
    worse than twice Netlib's own (normwise and mean), byte-identical reruns; the verifier
    verdict against Netlib is asserted for the planner's own choice (a K split of the packed
    kernels or k slices of the one-launch kernel, per the tuned table).
    """
    import cgemm3m

    m = n = 256
    k = 65536
    d = _tall_skinny_inputs(gpu)
    cfg = cgemm3m.config(k_split=parts)
    print("case tall-skinny parts=%d" % parts)
    sys.stdout.flush()
    st, out = run_gpu(gpu, "N", "N", m, n, k, ALPHA2, d["A"], m, d["B"], k, BETA2, d["C"], m, cfg, d["dA"], d["dB"])
    assert st == 0
    plan = gpu.last_plan(sync=True)
    print("plan", plan)
    assert plan["parts"] == parts if parts else plan["parts"] > 1
    dOut = to_dev(out)
    mine = cgemm3m.residual(gpu, "N", "N", m, n, k, ALPHA2, ptr(d["dA"]), m, ptr(d["dB"]), k, BETA2, ptr(d["dC0"]),
                            ptr(dOut), m)
    ref = d["netlib_err"]
    print("fp64 error: ours %s netlib %s" % (mine, ref))
    # the stated 1e-5, widened to twice Netlib's own normwise FP64 error where that is larger
    tol = max(1e-5, 2 * ref["normwise"])
    ok, worst = oracle.reltol_check(out, d["gold"], rtol=tol, atol_frac=tol)
    print("reltol at %.3g: worst %.3g" % (tol, worst))
    assert ok, worst
    assert mine["normwise"] <= 2 * ref["normwise"] and mine["mean_rel"] <= 2 * ref["mean_rel"]
    st2, out2 = run_gpu(gpu, "N", "N", m, n, k, ALPHA2, d["A"], m, d["B"], k, BETA2, d["C"], m, cfg, d["dA"],
                        d["dB"])
    same = out2.tobytes() == out.tobytes()
    assert same, "second run differs"
    if parts == 0:
        check_against_oracle(out, d["gold"], d["C"], m, n, m, "tall-skinny-auto")


@pytest.mark.parametrize("prec", ["bf16x3", "bf16x6"])
def test_tall_skinny_bf16_products_ksplit(gpu, prec):
    """256 x 256 x 65536 on the BF16 forms: the planner splits k over the separate-products
    kernel (too few 256 x 256 tiles for the CUs), and merge_kernel sums each product's
    partitions in index order. FP64 error no worse than twice Netlib's own (normwise; the
    bf16x3 products drop the lo*lo term, so its mean elementwise error is about 3x the FP32
    path's and is bounded at 4x Netlib's), byte-identical reruns."""
    import cgemm3m

    m = n = 256
    k = 65536
    d = _tall_skinny_inputs(gpu)
    cfg = cgemm3m.config(prec)
    print("case tall-skinny", prec)
    sys.stdout.flush()
    st, out = run_gpu(gpu, "N", "N", m, n, k, ALPHA2, d["A"], m, d["B"], k, BETA2, d["C"], m, cfg, d["dA"], d["dB"])
    assert st == 0
    plan = gpu.last_plan(sync=True)
    print("plan", plan)
    assert plan["path"] == 4 and plan["parts"] > 1
    dOut = to_dev(out)
    mine = cgemm3m.residual(gpu, "N", "N", m, n, k, ALPHA2, ptr(d["dA"]), m, ptr(d["dB"]), k, BETA2, ptr(d["dC0"]),
                            ptr(dOut), m)
    ref = d["netlib_err"]
    print("fp64 error: ours %s netlib %s" % (mine, ref))
    assert mine["normwise"] <= 2 * ref["normwise"] and mine["mean_rel"] <= 4 * ref["mean_rel"]
    st2, out2 = run_gpu(gpu, "N", "N", m, n, k, ALPHA2, d["A"], m, d["B"], k, BETA2, d["C"], m, cfg, d["dA"],
                        d["dB"])
    assert out2.tobytes() == out.tobytes(), "second run differs"


@pytest.mark.parametrize("identity", [1, 2])
@pytest.mark.parametrize("tile", [64, 128])
def test_identity_and_tile(gpu, identity, tile):
    import cgemm3m

    cfg = cgemm3m.config(tile=tile, identity=identity)
    _case(gpu, "C", "T", 300, 200, 333, ALPHA2, BETA2, 333 + 3, 200 + 5, 300 + 1, 20260500 + tile + identity,
          cfg=cfg, label="id%d-t%d" % (identity, tile))


@pytest.mark.parametrize("prec,variants", [("fp32", (1, 2, 3)), ("bf16x3", (1, 2, 4, 6)), ("bf16x6", (2, 4, 6)),
                                           ("fp16", (1, 2, 4, 6))])
@pytest.mark.parametrize("parts", [1, 3])
def test_kernel_variants_identical(gpu, prec, variants, parts):
    """The product kernels (four or eight waves through LDS, four waves loading fragments
    straight into registers, each real product on its own 256x256 tiles then combined) issue
    the same MFMAs per accumulator in the same order and combine the products the same way,
    so their outputs are byte-identical; the first also passes the oracle check."""
    import cgemm3m

    m, n, k = 1000, 700, 1100
    variants = [v for v in variants if not (v in (4, 6) and parts > 1)]  # no K split there
    outs = []
    for v in variants:
        cfg = cgemm3m.config(prec, tile=128, k_split=parts, variant=v)
        if v == variants[0] and prec == "fp32":
            out, _ = _case(gpu, "C", "N", m, n, k, ALPHA2, BETA2, k + 3, k + 1, m + 2, 20260550 + parts, cfg=cfg,
                           label="variant%d" % v)
        else:
            A, B, C = inputs("C", "N", m, n, k, k + 3, k + 1, m + 2, 20260550 + parts)
            st, out = run_gpu(gpu, "C", "N", m, n, k, ALPHA2, A, k + 3, B, k + 1, BETA2, C, m + 2, cfg)
            assert st == 0
        outs.append(out)
    for v, o in zip(variants[1:], outs[1:]):
        same = o.tobytes() == outs[0].tobytes()  # (no element diff: it can run for minutes)
        assert same, "variant %d differs from variant %d in %d outputs" % (
            v, variants[0], int(np.count_nonzero(o.view(np.uint64) != outs[0].view(np.uint64))))


RAW_BF16_CASES = [
    # (m, n, k, ops, parts): ragged edges, all op layouts, K partitions, a k tail
    (200, 250, 3000, "NN", 4), (256, 256, 4096, "TC", 0), (130, 70, 1000, "CN", 1), (64, 300, 2048, "NT", 3),
    (300, 200, 777, "TT", 2), (128, 128, 8192, "CC", 0),
]


@pytest.mark.parametrize("prec", ["bf16x3", "bf16x6"])
@pytest.mark.parametrize("case", RAW_BF16_CASES, ids=["%dx%dx%d-%s-p%d" % c for c in RAW_BF16_CASES])
def test_bf16_raw_products_identical(gpu, prec, case):
    """BF16 products read from the stored operands (variant 7) against the same products on
    planes packed first (variant 4), with the same K partitions: the plane values and BF16
    pieces are formed with the pack's arithmetic and every accumulator sees the same MFMAs in
    the same order, so the outputs are byte-identical; bf16x6's also passes the oracle check
    (bf16x3 drops the lo*lo terms and can miss the reference verifier's FP32 match ratio).
    One case scales a row block of A down to FP32 denormals (kept by both splits)."""
    import cgemm3m

    m, n, k, ops, parts = case
    ta, tb = ops[0], ops[1]
    lda = (m if ta == "N" else k) + 3
    ldb = (k if tb == "N" else n) + 1
    ldc = m + 2
    seed = 20261204 + m + k
    A, B, C = inputs(ta, tb, m, n, k, lda, ldb, ldc, seed)
    denormal = ops == "TT"
    if denormal:
        # op(A) rows 0..39 (A stored k x m, so its first 40 columns) scaled to FP32 denormals:
        # both splits keep them; BF16 pieces of denormals carry few bits, so no oracle check
        A = A.copy()
        A[: 40 * lda] *= 1e-39
    outs = []
    for v in (7, 4):  # the raw form first: its plan's partitions (parts 0: its own choice) pin 4's
        cfg = cgemm3m.config(prec, k_split=parts, variant=v)
        print("case raw-bf16", prec, case, "variant", v, "parts", parts)
        sys.stdout.flush()
        if v == 4 and not denormal and prec == "bf16x6":  # bf16x3 misses the verifier at times
            out, _ = _case(gpu, ta, tb, m, n, k, ALPHA2, BETA2, lda, ldb, ldc, seed, cfg=cfg,
                           label="raw-%s-v%d" % (prec, v))
        elif v == 4:
            st, out = run_gpu(gpu, ta, tb, m, n, k, ALPHA2, A, lda, B, ldb, BETA2, C, ldc, cfg)
            assert st == 0
        else:
            st, out = run_gpu(gpu, ta, tb, m, n, k, ALPHA2, A, lda, B, ldb, BETA2, C, ldc, cfg)
            assert st == 0
            plan = gpu.last_plan(sync=True)
            assert plan["path"] == 4 and plan["tile"] == 128, plan
            parts = plan["parts"]
        outs.append(out)
    diff = int(np.count_nonzero(outs[1].view(np.uint64) != outs[0].view(np.uint64)))
    assert diff == 0, "%d outputs differ between the raw and packed products" % diff


@pytest.mark.parametrize("parts", [1, 3])
def test_tile_sizes_identical(gpu, parts):
    """64 x 64 tiles through LDS (four waves of 32 x 32) and 128 x 128 tiles loading
    fragments straight into registers (four waves of 64 x 64) issue the same MFMAs per
    accumulator in the same k order, so the output bytes do not depend on the tile size (the
    grouped batch relies on it: it computes every member on 64 x 64 wave tiles)."""
    import cgemm3m

    m, n, k = 300, 200, 700
    A, B, C = inputs("T", "C", m, n, k, k + 1, n + 2, m + 3, 20260580 + parts)
    outs = []
    for tile, variant in ((64, 1), (128, 3), (128, 1)):
        cfg = cgemm3m.config(tile=tile, k_split=parts, variant=variant)
        st, out = run_gpu(gpu, "T", "C", m, n, k, ALPHA2, A, k + 1, B, n + 2, BETA2, C, m + 3, cfg)
        assert st == 0
        outs.append(out)
    for i in (1, 2):
        same = outs[i].tobytes() == outs[0].tobytes()
        assert same, "form %d differs from the 64 x 64 tile in %d outputs" % (
            i, int(np.count_nonzero(outs[i].view(np.uint64) != outs[0].view(np.uint64))))


def _no_neg_zero(x):
    """The bytes of x with -0.0 read as +0.0: an exactly cancelling sum may carry either sign
    of zero depending on its order, which the comparison leaves open."""
    return np.asarray(x).view(np.float32) + np.float32(0.0)


def _int_fill(rows, cols, ld, rng, lim=8):
    """Column-major complex64 storage with integer real and imaginary parts in [-lim, lim]."""
    x = rng.integers(-lim, lim + 1, size=(2 * max(ld, 1) * max(cols, 0),)).astype(np.float32)
    return x.view(np.complex64).copy()


INT_CASES = [
    # (label, m, n, k, ops, config kwargs)
    ("one-launch", 100, 70, 200, "NN", dict()),
    ("one-launch-1wave", 90, 33, 300, "CT", dict(variant=5, tile=16, k_split=1)),
    ("one-launch-sliced", 40, 40, 512, "TN", dict(variant=5, k_split=4)),
    ("tile64-lds", 300, 200, 333, "NC", dict(tile=64, k_split=1)),
    ("tile128-lds", 300, 200, 333, "TT", dict(tile=128, k_split=1, variant=1)),
    ("tile128-direct", 300, 200, 333, "CN", dict(tile=128, k_split=1, variant=3)),
    ("k-split3", 260, 130, 512, "NN", dict(tile=128, k_split=3)),
    ("four-products", 300, 200, 333, "CC", dict(identity=3)),
    ("standard-identity", 300, 200, 333, "NT", dict(identity=2, tile=64)),
    ("bf16x3", 300, 260, 512, "NN", dict(precision="bf16x3")),
    ("bf16x6", 300, 260, 512, "CT", dict(precision="bf16x6")),
    ("fp16", 300, 260, 512, "TC", dict(precision="fp16")),
    ("bf16x3-products", 2048, 1024, 512, "NN", dict(precision="bf16x3", variant=4)),
    ("fp16-products", 2048, 1024, 512, "CN", dict(precision="fp16", variant=4)),
    ("bf16x3-products-ksplit5", 300, 260, 2048, "CN", dict(precision="bf16x3", variant=4, k_split=5)),
    ("bf16x6-products-ksplit3", 260, 300, 1024, "TC", dict(precision="bf16x6", variant=4, k_split=3)),
    ("bf16x3-tall-skinny-auto", 200, 256, 8192, "NT", dict(precision="bf16x3")),
    ("bf16x3-raw-ksplit4", 200, 250, 3000, "CT", dict(precision="bf16x3", variant=7, k_split=4)),
    ("bf16x6-raw", 130, 70, 1000, "TN", dict(precision="bf16x6", variant=7)),
]


@pytest.mark.parametrize("case", INT_CASES, ids=[c[0] for c in INT_CASES])
def test_integer_inputs_bitwise(gpu, case):
    """Integer real and imaginary parts in [-8, 8] with |partial sums| far below 2^24: every
    product and sum of every path is exact in FP32 (the bf16 hi/lo split leaves lo = 0, the
    FP16 block scales are powers of two), and so is the reference BLAS's order, so each path
    must equal the oracle bit for bit. alpha and beta are small Gaussian integers."""
    import cgemm3m

    label, m, n, k, ops, kw = case
    ta, tb = ops[0], ops[1]
    rng = np.random.default_rng(zlib.crc32(label.encode()))
    lda = (m if ta == "N" else k) + 1
    ldb = (k if tb == "N" else n) + 2
    ldc = m + 3
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    A, B, C = _int_fill(ar, ac, lda, rng), _int_fill(br, bc, ldb, rng), _int_fill(m, n, ldc, rng)
    alpha, beta = complex(2, -1), complex(1, 1)
    prec = kw.pop("precision", "fp32")
    cfg = cgemm3m.config(prec, **kw)
    print("case", label, ops, m, n, k, prec, kw)
    sys.stdout.flush()
    st, out = run_gpu(gpu, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, cfg)
    assert st == 0
    gold = oracle.cgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    same = _no_neg_zero(out).tobytes() == _no_neg_zero(gold).tobytes()
    assert same, "%s: %d outputs differ from the oracle" % (
        label, int(np.count_nonzero(out.view(np.uint64) != gold.view(np.uint64))))


@pytest.mark.parametrize("shapes", [
    # quad-kernel work dominates: the members the table gives to the one-launch kernel join it
    [(300, 200, 333), (40, 48, 512), (512, 512, 256), (100, 30, 64), (256, 256, 1024)],
    # one-launch work dominates: both kernels run in the one call
    [(400, 400, 64)] + [(64, 64, 1024)] * 8 + [(100, 30, 64)],
], ids=["quad-heavy", "one-launch-heavy"])
def test_integer_inputs_bitwise_grouped(gpu, shapes):
    """The same exactness through one grouped call (quad-kernel members read as stored,
    one-launch members, K-split members per the tuned table)."""
    import torch

    import cgemm3m

    rng = np.random.default_rng(99)
    probs, keep, gold = [], [], []
    for i, (m, n, k) in enumerate(shapes):
        ta, tb = "NTC"[i % 3], "NTC"[(i + 1) % 3]
        lda = (m if ta == "N" else k) + 1
        ldb = (k if tb == "N" else n) + 1
        A = _int_fill(*((m, k) if ta == "N" else (k, m)), lda, rng)
        B = _int_fill(*((k, n) if tb == "N" else (n, k)), ldb, rng)
        C = _int_fill(m, n, m, rng)
        dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
        keep += [dA, dB, dC]
        probs.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=complex(1, 2), A=ptr(dA), lda=lda, B=ptr(dB), ldb=ldb,
                          beta=complex(-1, 0), C=ptr(dC), ldc=m))
        gold.append(oracle.cgemm(ta, tb, m, n, k, complex(1, 2), A, lda, B, ldb, complex(-1, 0), C, m))
    cgemm3m.grouped(gpu, probs)
    torch.cuda.synchronize()
    for i, g in enumerate(gold):
        out = keep[3 * i + 2].cpu().numpy()
        same = _no_neg_zero(out).tobytes() == _no_neg_zero(g).tobytes()
        assert same, "member %d %s: %d outputs differ" % (i, shapes[i], int(np.count_nonzero(out != g)))


@pytest.mark.parametrize("seed", range(6))
def test_small_ragged(gpu, seed):
    rng = np.random.default_rng(seed)
    ta, tb = rng.choice(list("NTC")), rng.choice(list("NTC"))
    m, n, k = (int(x) for x in rng.integers(1, 70, 3))
    lda = (m if ta == "N" else k) + int(rng.integers(0, 4))
    ldb = (k if tb == "N" else n) + int(rng.integers(0, 4))
    ldc = m + int(rng.integers(0, 4))
    _case(gpu, str(ta), str(tb), m, n, k, ALPHA2, BETA2, lda, ldb, ldc, 20260600 + seed, label="small")


@pytest.mark.parametrize("kw", [
    dict(m=1, n=1, k=1, alpha=1, beta=0),                              # m1n1k1
    dict(m=32, n=32, k=32, ta="C", tb="C", alpha=complex(100, 50), beta=0),  # large_alpha
    dict(m=32, n=32, k=32, ta="T", tb="C", alpha=1j, beta=-1j),        # pure_imag_scalars
    dict(m=16, n=16, k=0, alpha=1, beta=0),                            # k0_beta0_C_zeroed
    dict(m=16, n=16, k=0, alpha=1, beta=1),                            # k0_beta1_C_unchanged
    dict(m=16, n=16, k=16, alpha=0, beta=complex(0.5, 0.5)),           # alpha0_beta_half
    dict(m=16, n=16, k=16, alpha=0, beta=1),                           # alpha0_beta1
])
def test_edge_rows(gpu, kw):
    kw = dict(kw)
    ta, tb = kw.pop("ta", "N"), kw.pop("tb", "N")
    m, n, k = kw["m"], kw["n"], kw["k"]
    lda = max(1, m if ta == "N" else k)
    ldb = max(1, k if tb == "N" else n)
    _case(gpu, ta, tb, m, n, k, kw["alpha"], kw["beta"], lda, ldb, max(1, m), 20260700 + m + k, label="edge")


def test_beta0_ignores_nan_c(gpu):
    """beta0_c_nan: C filled with NaN is never read when beta == 0."""
    m = n = k = 32
    A, B, _ = inputs("N", "N", m, n, k, m, k, m, 20260800)
    C = np.full(m * n, complex(np.nan, np.nan), dtype=np.complex64)
    st, out = run_gpu(gpu, "N", "N", m, n, k, 1, A, m, B, k, 0, C, m)
    assert st == 0 and np.isfinite(out.view(np.float32)).all()
    gold = oracle.cgemm("N", "N", m, n, k, 1, A, m, B, k, 0, C, m)
    check_against_oracle(out, gold, C, m, n, m, "beta0_nan")


def test_zero_product_allows_null_inputs(gpu):
    """Reference KAT ZeroProductAllowsNullInputs (cgemm3m_test.cpp:189-210)."""
    for zero_k in (True, False):
        k = 0 if zero_k else 7
        alpha = 1.0 if zero_k else 0.0
        for beta in (0, 1, complex(0.5, -0.25)):
            inp = np.nan if beta == 0 else 7.0
            init = np.full(10, complex(inp, -3), dtype=np.complex64)
            st, out = run_gpu(gpu, "N", "N", 3, 2, k, alpha, None, 3, None, 7, beta, init, 5)
            assert st == 0
            gold = oracle.cgemm("N", "N", 3, 2, k, alpha, None, 3, None, 7, beta, init, 5)
            ok, stats = oracle.verify(out, gold, init, 3, 2, 5)
            assert ok, stats


def test_empty_output_preserves_sentinel(gpu):
    sentinel = np.full(12, complex(7, -3), dtype=np.complex64)
    for m, n in ((0, 3), (3, 0)):
        st, out = run_gpu(gpu, "N", "N", m, n, 2, 1, None, 4, None, 4, 0, sentinel, 4)
        same = st == 0 and out.tobytes() == sentinel.tobytes()
        assert same


def test_segment_cancellation(gpu):
    """Reference KAT SegmentCancellation (cgemm3m_test.cpp:212-240): C(0,0) must be r within
    0.01 and the whole output must pass the verifier against the exact result."""
    size = 2048
    a = np.zeros(size * size, np.complex64)
    b = np.zeros(size * size, np.complex64)
    for l in range(40 * 33):
        a[l * size] = 1
        b[l] = 4096.1 if l < 20 * 33 else -4096.1
    a[(size - 1) * size] = 1
    init = np.zeros(size * size, np.complex64)
    dA = to_dev(a)
    for r in (0.0, 1.0):
        b[size - 1] = r
        expected = init.copy()
        expected[0] = r
        st, out = run_gpu(gpu, "N", "N", size, size, size, 1, a, size, b, size, 0, init, size, dA=dA)
        assert st == 0
        print("segment cancellation r=%g got %r plan %s" % (r, out[0], gpu.last_plan(sync=True)))
        assert abs(float(out[0].real) - r) <= 0.01 and abs(float(out[0].imag)) <= 0.01
        ok, stats = oracle.verify(out, expected, init, size, size, size)
        assert ok, stats


def test_residual_matches_cpu_fp64(gpu):
    import cgemm3m

    m, n, k = 70, 50, 90
    A, B, C = inputs("T", "C", m, n, k, k, n, m, 20260900)
    dA, dB, dC, dC0 = to_dev(A), to_dev(B), to_dev(C), to_dev(C)
    cgemm3m.cgemm3m(gpu, "T", "C", m, n, k, ALPHA2, ptr(dA), k, ptr(dB), n, BETA2, ptr(dC), m)
    res = cgemm3m.residual(gpu, "T", "C", m, n, k, ALPHA2, ptr(dA), k, ptr(dB), n, BETA2, ptr(dC0), ptr(dC), m)
    z = oracle.zgemm("T", "C", m, n, k, ALPHA2, A, k, B, n, BETA2, C, m)
    ref = oracle.errors_vs(dC.cpu().numpy(), z)
    assert abs(res["max_rel"] - ref["max_rel"]) <= 1e-6 * max(1.0, ref["max_rel"])
    assert abs(res["normwise"] - ref["normwise"]) <= 1e-9
    assert res["normwise"] < 1e-6


def test_correction_fixes_flagged_and_counts(gpu):
    """Forced correction on a cancellation-heavy input: flagged outputs are recomputed in FP64
    and land on the FP64 result; forcing it off leaves the FP32 error."""
    import cgemm3m

    m, n, k = 64, 64, 4096
    A, B, C = inputs("N", "N", m, n, k, m, k, m, 20261000)
    # make column 0 of B nearly orthogonal to row 0 of A: b[:,0] = t - proj
    Ad = oracle.op_matrix(A, m, k, m, "N")
    Bd = oracle.op_matrix(B, k, n, k, "N")
    x = Ad[0, :]
    v = Bd[:, 0] * 1000.0
    v = v - (x @ v) / (x @ np.conj(x).T).real * np.conj(x)
    B2 = B.copy()
    B2[:k] = v.astype(np.complex64)
    dA, dB = to_dev(A), to_dev(B2)
    z = oracle.zgemm("N", "N", m, n, k, 1, A, m, B2, k, 0, C, m)
    outs = {}
    for corr in (1, -1):
        st, out = run_gpu(gpu, "N", "N", m, n, k, 1, A, m, B2, k, 0, C, m, cgemm3m.config(correction=corr), dA, dB)
        assert st == 0
        outs[corr] = (out, gpu.last_plan(sync=True))
        print("correction", corr, outs[corr][1], out[0], z[0])
    on, plan_on = outs[1]
    assert plan_on["flagged"] >= 1
    assert abs(complex(on[0]) - z[0]) <= 1e-6 * abs(z[0]) + 1e-30


@pytest.mark.parametrize("pin", [dict(), dict(variant=5, k_split=1), dict(variant=5, k_split=4),
                                 dict(tile=64, k_split=1), dict(tile=128, k_split=2)])
def test_correction_in_reference_blas_order(gpu, pin):
    """correction = 2: flagged outputs are recomputed in the reference BLAS's sequential FP32
    order, so they are bit-identical to the Netlib-order oracle; the rest meet the mixed
    tolerance.
    Planner's choice, the one-launch kernel unsliced and with k in four slices (in-kernel
    recompute by the tile's last workgroup), the packed tile path (recompute kernel) and the
    K-split path (flags raised by the merge)."""
    import cgemm3m

    m, n, k = 64, 64, 4096
    A, B, C = inputs("N", "N", m, n, k, m, k, m, 20261000)
    Ad = oracle.op_matrix(A, m, k, m, "N")
    Bd = oracle.op_matrix(B, k, n, k, "N")
    x = Ad[0, :]
    v = Bd[:, 0] * 1000.0
    v = v - (x @ v) / (x @ np.conj(x).T).real * np.conj(x)
    B2 = B.copy()
    B2[:k] = v.astype(np.complex64)
    for beta in (0, BETA2):
        cfg = cgemm3m.config(correction=2, **pin)
        st, out = run_gpu(gpu, "N", "N", m, n, k, ALPHA2, A, m, B2, k, beta, C, m, cfg)
        assert st == 0
        plan = gpu.last_plan(sync=True)
        print("plan", plan)
        assert plan["correction"] == 2 and plan["flagged"] >= 1
        gold = oracle.cgemm("N", "N", m, n, k, ALPHA2, A, m, B2, k, beta, C, m)
        assert out[0].tobytes() == gold[0].tobytes(), (out[0], gold[0])
        # column 0 is scaled by 1e3, so its other outputs are ~1e4-1e5 in size: the mixed
        # tolerance applies (the verifier's 32-ulp cap is made for the reference's inputs)
        ok, worst = oracle.reltol_check(region(out, m, n, m), region(gold, m, n, m))
        assert ok, worst


def test_golden_fixtures(gpu):
    """The committed fixtures (tests/golden): every case through the C ABI against its stored
    Netlib-order result (reltol, verifier, padding rows untouched, NaN padding kept)."""
    import json
    import os

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    data = np.load(os.path.join(here, "golden.npz"))
    with open(os.path.join(here, "golden.json")) as f:
        meta = json.load(f)
    for i, c in enumerate(meta):
        print("case golden %d %s" % (i, c))
        sys.stdout.flush()
        A, B, C = data["%d_A" % i], data["%d_B" % i], data["%d_C" % i]
        st, out = run_gpu(gpu, c["ta"], c["tb"], c["m"], c["n"], c["k"], complex(*c["alpha"]), A, c["lda"], B,
                          c["ldb"], complex(*c["beta"]), C, c["ldc"])
        assert st == 0
        check_against_oracle(out, data["%d_out32" % i], C, c["m"], c["n"], c["ldc"], "golden%d" % i)


@pytest.mark.parametrize("ta,tb", list(itertools.product("NTC", "NTC")))
@pytest.mark.parametrize("identity", [1, 2])
def test_small_kernel_ops(gpu, ta, tb, identity):
    """The one-launch small-problem kernel (variant 5: reads A and B directly, 16x16 output per
    wave) on ragged shapes with padded leading dimensions, every op pair, both identities."""
    import cgemm3m

    m, n, k = 100, 70, 91
    lda = (m if ta == "N" else k) + 3
    ldb = (k if tb == "N" else n) + 2
    cfg = cgemm3m.config(variant=5, identity=identity)
    out, plan = _case(gpu, ta, tb, m, n, k, ALPHA2, BETA2, lda, ldb, m + 5,
                      20260800 + 9 * identity + 3 * "NTC".index(ta) + "NTC".index(tb), cfg=cfg, label="small")
    assert plan["path"] == 5


@pytest.mark.parametrize("slices", [2, 3, 7, 16])
def test_small_kernel_k_slices(gpu, slices):
    """The small-problem kernel with its k range split over `slices` workgroups per tile (the
    last to arrive adds the slices in slice order): a ragged tall-skinny shape with padded
    leading dimensions, beta != 0, with the correction (k >= 1024: energies summed over the
    slices, flagged outputs recomputed) and without it; two calls in a row (the arrival
    counters are left zero for the next launch), byte-identical reruns."""
    import cgemm3m

    for (m, n, k, ta, tb), correction in (((181, 37, 1500, "T", "N"), 1), ((90, 70, 700, "N", "C"), -1)):
        lda = (m if ta == "N" else k) + 3
        ldb = (k if tb == "N" else n) + 2
        cfg = cgemm3m.config(variant=5, k_split=slices, correction=correction)
        out, plan = _case(gpu, ta, tb, m, n, k, ALPHA2, BETA2, lda, ldb, m + 1, 20260900 + slices + k, cfg=cfg,
                          label="small-sliced%d" % slices)
        assert plan["path"] == 5 and plan["parts"] == slices, plan
        assert (plan["flagged"] >= 0) == (correction == 1), plan


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "C"), ("C", "T")])
def test_small_kernel_single_wave(gpu, ta, tb):
    """The one-launch kernel's single-wave form (tile 16: one wave per 16x16 tile takes every k
    chunk, 64-thread workgroups): ragged short-k and long-k shapes with padded leading
    dimensions, unsliced and with k in three slices, the correction on (k >= 1024) and off."""
    import cgemm3m

    for (m, n, k), slices in (((100, 70, 40), 1), ((181, 37, 1500), 1), ((181, 37, 1500), 3)):
        lda = (m if ta == "N" else k) + 3
        ldb = (k if tb == "N" else n) + 2
        cfg = cgemm3m.config(variant=5, tile=16, k_split=slices)
        out, plan = _case(gpu, ta, tb, m, n, k, ALPHA2, BETA2, lda, ldb, m + 1,
                          20261600 + 3 * "NTC".index(ta) + "NTC".index(tb) + k + slices, cfg=cfg,
                          label="small-1wave-s%d" % slices)
        assert plan["path"] == 5 and plan["parts"] == slices, plan


def test_small_kernel_sliced_calls_back_to_back(gpu):
    """k-sliced one-launch calls share the handle's arrival counters, which each launch
    leaves at zero: calls of different tile counts queued back to back (no synchronisation
    between them) all meet the oracle, and repeating the first gives its bytes again."""
    import torch

    import cgemm3m

    shapes = [(70, 50, 900), (300, 200, 640), (70, 50, 900), (33, 400, 1200)]
    runs = []
    for idx, (m, n, k) in enumerate(shapes):
        A, B, C = inputs("N", "C", m, n, k, m, n, m, 20262000 + m + n + k)
        dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
        cfg = cgemm3m.config(variant=5, k_split=3 + idx % 2)
        st = cgemm3m.cgemm3m(gpu, "N", "C", m, n, k, ALPHA2, ptr(dA), m, ptr(dB), n, BETA2, ptr(dC), m, cfg)
        assert st == 0
        runs.append((m, n, k, A, B, C, dA, dB, dC))
    torch.cuda.synchronize()
    outs = []
    for m, n, k, A, B, C, dA, dB, dC in runs:
        out = dC.cpu().numpy()
        gold = oracle.cgemm("N", "C", m, n, k, ALPHA2, A, m, B, n, BETA2, C, m)
        check_against_oracle(out, gold, C, m, n, m, "sliced-back-to-back")
        outs.append(out)
    same = outs[0].tobytes() == outs[2].tobytes()
    assert same


def test_small_kernel_auto_slices(gpu):
    """Few 16x16 tiles and a long k: the planner slices k on its own (about two workgroups
    per CU)."""
    m, n, k = 464, 21, 1663
    out, plan = _case(gpu, "N", "T", m, n, k, ALPHA2, BETA2, m, n, m, 20260950, label="small-auto-sliced")
    assert plan["path"] == 5 and plan["parts"] > 1, plan


@pytest.mark.parametrize("shape", [(200, 150, 300), (1, 1, 1), (33, 17, 1000), (64, 300, 5)])
def test_small_kernel_auto(gpu, shape):
    """Shapes the planner sends to the small-problem kernel by default (correction off,
    beta == 0 included)."""
    m, n, k = shape
    for beta in (BETA2, 0):
        out, plan = _case(gpu, "N", "C", m, n, k, ALPHA2, beta, m, n, m, 20260850 + m + n + k, label="small-auto")
        assert plan["path"] == 5
