"""Special values and cancellation on every path, against the Netlib-order oracle's verdict.

The reference's FL rows (ops-blas/test/gemm/cgemm3m/arch35/cgemm3m_test.csv:257-274) fill A, B
or C (with beta = (0, 1)) with zeros, alternating signs, extreme magnitudes, Inf and NaN and
judge the result with the verifier's non-finite rules (cgemm3m_verify.h:31-38: NaN and Inf
are equivalent, a finite value against a non-finite one fails, opposite infinities fail).
The 3M planes can overflow where the reference BLAS does not (A = (3e38, 3e38) times
B = (1, 0)); every epilogue recomputes such outputs in the reference BLAS's order, so the
finite/non-finite pattern must match the oracle's on the one-launch kernel, the packed fused
tile, the K split, strided batches and grouped batches, with the rounding correction off and
on. The reference's fill generator (fill.h) is not part of its snapshot, so the fills here are
this project's own: the same categories, seeded.

SegmentCancellation (cgemm3m_test.cpp:212-240) as a member of a strided and of a grouped
batch: members get the single call's rounding correction (k >= 1024), so the member returns
the exact answer there too.
"""
import sys

import numpy as np
import pytest

import oracle
from helpers import ptr, region, to_dev

pytestmark = pytest.mark.gpu

OPS = [a + b for a in "NTC" for b in "NTC"]
BIG = np.float32(3.0e38)


def _store(opX, t, rows_st, cols_st, ld):
    """Column-major storage (ld x cols_st) of the stored matrix whose op(.) is opX."""
    S = opX if t == "N" else (opX.T if t == "T" else opX.T.conj())
    assert S.shape == (rows_st, cols_st)
    out = np.zeros((cols_st, ld), np.complex64)
    out[:, :rows_st] = S.T
    return out.ravel()


def _logical(kind, rows, cols, seed, unique_axis):
    """rows x cols complex64 op(X) for one fill category. unique_axis (0: each row, 1: each
    column) gets at most one extreme entry, so a sum of products holds at most one term near
    the overflow threshold and no partial sum of the reference's order overflows before
    cancelling."""
    rng = np.random.default_rng(seed)
    base = (rng.uniform(-1, 1, (rows, cols)) + 1j * rng.uniform(-1, 1, (rows, cols))).astype(np.complex64)
    if kind == "rand":
        return base
    if kind == "zero":
        return np.zeros((rows, cols), np.complex64)
    if kind == "alter":  # alternating signs along both axes
        sign = np.where((np.arange(rows)[:, None] + np.arange(cols)[None, :]) % 2 == 0, 1.0, -1.0)
        return (np.abs(base.real) * sign + 1j * np.abs(base.imag) * sign).astype(np.complex64)
    if kind in ("inf_all", "nan_all"):
        v = np.float32(np.inf if kind == "inf_all" else np.nan)
        return np.full((rows, cols), complex(v, v), np.complex64)
    if kind in ("inf_some", "nan_some"):
        x = base.copy()
        v = np.float32(np.inf if kind == "inf_some" else np.nan)
        hit = rng.random((rows, cols)) < 0.03
        part = rng.integers(0, 3, (rows, cols))  # real, imag or both
        sgn = np.where(rng.random((rows, cols)) < 0.5, 1, -1).astype(np.float32)
        re, im = x.real.copy(), x.imag.copy()
        with np.errstate(invalid="ignore"):
            re[hit & (part != 1)] = (v * sgn)[hit & (part != 1)]
            im[hit & (part != 0)] = (v * sgn)[hit & (part != 0)]
            return (re + 1j * im).astype(np.complex64)
    if kind == "extreme":
        # one entry per row (or column) near the overflow threshold with random signs, and
        # subnormal and tiny normal entries elsewhere
        x = base.copy()
        re, im = x.real.copy(), x.imag.copy()
        n_lines = rows if unique_axis == 0 else cols
        for t in range(n_lines):
            if rng.random() < 0.5:
                continue
            pos = int(rng.integers(0, cols if unique_axis == 0 else rows))
            r, c = (t, pos) if unique_axis == 0 else (pos, t)
            re[r, c] = np.float32(rng.uniform(1.0e38, 3.4e38)) * np.float32(rng.choice([-1, 1]))
            im[r, c] = np.float32(rng.uniform(1.0e38, 3.4e38)) * np.float32(rng.choice([-1, 1]))
        sub = rng.random((rows, cols)) < 0.1
        re[sub] = (rng.uniform(-1, 1, sub.sum()) * 1e-39).astype(np.float32)  # subnormal
        tiny = (rng.random((rows, cols)) < 0.1) & ~sub
        im[tiny] = (rng.uniform(-1, 1, tiny.sum()) * 1e-30).astype(np.float32)
        return (re + 1j * im).astype(np.complex64)
    raise ValueError(kind)


def _problem(ta, tb, m, n, k, fill_a, fill_b, fill_c, seed, pad=(0, 0, 0)):
    lda = (m if ta == "N" else k) + pad[0]
    ldb = (k if tb == "N" else n) + pad[1]
    ldc = m + pad[2]
    opA = _logical(fill_a, m, k, seed, 0)
    opB = _logical(fill_b, k, n, seed + 1, 1)
    A = _store(opA, ta, *((m, k) if ta == "N" else (k, m)), lda)
    B = _store(opB, tb, *((k, n) if tb == "N" else (n, k)), ldb)
    C = _store(_logical(fill_c, m, n, seed + 2, 0), "N", m, n, ldc)
    return A, lda, B, ldb, C, ldc, opA, opB


def _nonfinite_agree(out, gold, m, n, ldc):
    """The verifier's non-finite rule on both components: the count of failures."""
    o = region(out, m, n, ldc)
    g = region(gold, m, n, ldc)
    bad = 0
    for part in (np.real, np.imag):
        _, st = oracle.verify_component(part(o).ravel(), part(g).ravel())
        bad += int(st[3])
    return bad


def _bound_check(out, gold, opA, opB, C0, alpha, beta, m, n, k, ldc, label):
    """Finite outputs within an elementwise GEMM error bound: |x - g| <= 8 (k + 8) 2^-24
    (|alpha| |op(A)| |op(B)| + |beta| |C|) per component (both the 3M result and the reference
    BLAS's are within (k + 8) u of that of the exact value; a dropped term, a transposed
    operand or a wrong sign breaks it by orders of magnitude)."""
    o = region(out, m, n, ldc).astype(np.complex128)
    g = region(gold, m, n, ldc).astype(np.complex128)
    with np.errstate(over="ignore", invalid="ignore"):
        mag = abs(alpha) * (np.abs(opA.astype(np.complex128)) @ np.abs(opB.astype(np.complex128)))
        mag = mag + abs(beta) * np.abs(region(C0, m, n, ldc).astype(np.complex128))
        bound = 8.0 * (k + 8) * 2.0 ** -24 * mag
        for part in (np.real, np.imag):
            fin = np.isfinite(part(o)) & np.isfinite(part(g)) & np.isfinite(bound)
            err = np.abs(part(o) - part(g))
            worst = np.max(np.where(fin, err - bound, -np.inf)) if fin.any() else -np.inf
            assert worst <= 0, "%s: finite output beyond the error bound by %g" % (label, worst)


def _run(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, cfg=None):
    import torch

    import cgemm3m

    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    st = cgemm3m.cgemm3m(h, ta, tb, m, n, k, alpha, ptr(dA), lda, ptr(dB), ldb, beta, ptr(dC), ldc, cfg,
                         check=False)
    torch.cuda.synchronize()
    assert st == 0, st
    return dC.cpu().numpy(), h.last_plan()


# fill categories for (A, B, C); C's special fills run with beta = (0, 1) as in the reference
FILLS = [("zero", "rand", "rand"), ("rand", "zero", "rand"), ("alter", "rand", "rand"), ("rand", "alter", "rand"),
         ("inf_some", "rand", "rand"), ("rand", "inf_some", "rand"), ("nan_some", "rand", "rand"),
         ("rand", "nan_some", "rand"), ("inf_all", "rand", "rand"), ("rand", "nan_all", "rand"),
         ("rand", "rand", "inf_some"), ("rand", "rand", "nan_some"), ("rand", "rand", "inf_all"),
         ("extreme", "rand", "rand"), ("rand", "extreme", "rand"), ("rand", "rand", "extreme")]

# forced paths besides the planner's own: packed fused tile (64 and 128), K split, the
# one-launch kernel with k slices, and the rounding correction forced on
PATHS = [dict(), dict(tile=64, k_split=1), dict(tile=128, k_split=1, variant=3), dict(tile=64, k_split=3),
         dict(variant=5, k_split=3), dict(tile=64, k_split=1, correction=1), dict(variant=5, correction=2)]


@pytest.mark.parametrize("fills", FILLS, ids=["-".join(f) for f in FILLS])
def test_special_values(gpu, fills):
    import cgemm3m

    fa, fb, fc = fills
    beta = complex(0, 1) if fc != "rand" else complex(0.5, -0.25)
    alpha = complex(1, 0) if fc != "rand" else complex(0.7, -0.3)
    runs = [(ops, 32, 32, 32, None) for ops in OPS]
    runs += [(ops, 300, 200, 333, None) for ops in OPS]
    runs += [(ops, 300, 200, 333, cfg) for ops in ("NN", "TC", "CT") for cfg in PATHS[1:]]
    for idx, (ops, m, n, k, pin) in enumerate(runs):
        ta, tb = ops
        label = "%s %s %dx%dx%d %s" % ("-".join(fills), ops, m, n, k, pin)
        print("case", label)
        sys.stdout.flush()
        A, lda, B, ldb, C, ldc, opA, opB = _problem(ta, tb, m, n, k, fa, fb, fc, 9100 + 7 * idx, pad=(1, 2, 3))
        cfg = cgemm3m.config(**pin) if pin else None
        out, plan = _run(gpu, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, cfg)
        gold = oracle.cgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
        bad = _nonfinite_agree(out, gold, m, n, ldc)
        assert bad == 0, "%s: %d outputs finite against non-finite (or opposite infinities); plan %s" % (
            label, bad, plan)
        pad = np.asarray(out).reshape(n, ldc)[:, m:].tobytes() == np.asarray(C).reshape(n, ldc)[:, m:].tobytes()
        assert pad, "%s: ldc padding written" % label
        if fa == "extreme" or fb == "extreme" or fc == "extreme":
            _bound_check(out, gold, opA, opB, C, alpha, beta, m, n, k, ldc, label)
        else:
            ok, stats = oracle.verify(out, gold, C, m, n, ldc)
            assert ok, (label, plan, stats)
        out2, _ = _run(gpu, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, cfg)
        assert out2.tobytes() == out.tobytes(), "%s: rerun differs" % label


def _check_member(out, gold, C, opA, opB, alpha, beta, m, n, k, ldc, fills, label):
    bad = _nonfinite_agree(out, gold, m, n, ldc)
    assert bad == 0, "%s: %d outputs finite against non-finite" % (label, bad)
    if "extreme" in fills:
        _bound_check(out, gold, opA, opB, C, alpha, beta, m, n, k, ldc, label)
    else:
        ok, stats = oracle.verify(out, gold, C, m, n, ldc)
        assert ok, (label, stats)


BATCH_FILLS = [("inf_some", "rand", "rand"), ("rand", "nan_some", "rand"), ("rand", "rand", "inf_some"),
               ("extreme", "rand", "rand"), ("rand", "extreme", "rand"), ("alter", "zero", "rand")]


@pytest.mark.parametrize("ops", ["NN", "CT", "TC"])
def test_special_values_strided(gpu, ops):
    """A strided batch of four members, each with its own fill (special values in A, B or C,
    extreme magnitudes), k = 333 (correction off) and k = 1100 (on)."""
    import torch

    import cgemm3m

    ta, tb = ops
    for m, n, k in ((300, 200, 333), (96, 80, 1100)):
        fills = [BATCH_FILLS[(i + (k > 1000)) % len(BATCH_FILLS)] for i in range(4)]
        beta = complex(0, 1)
        alpha = complex(1, 0)
        parts = [_problem(ta, tb, m, n, k, *f, 9600 + 11 * i) for i, f in enumerate(fills)]
        lda, ldb, ldc = parts[0][1], parts[0][3], parts[0][5]
        sa, sb, sc = parts[0][0].size, parts[0][2].size, parts[0][4].size
        A = np.concatenate([p[0] for p in parts])
        B = np.concatenate([p[2] for p in parts])
        C = np.concatenate([p[4] for p in parts])
        print("strided special", ops, m, n, k, fills)
        sys.stdout.flush()
        dA, dB = to_dev(A), to_dev(B)
        outs = []
        for _ in range(2):
            dC = to_dev(C)
            cgemm3m.strided_batched(gpu, ta, tb, m, n, k, alpha, ptr(dA), lda, sa, ptr(dB), ldb, sb, beta, ptr(dC),
                                    ldc, sc, 4)
            torch.cuda.synchronize()
            outs.append(dC.cpu().numpy())
        assert outs[0].tobytes() == outs[1].tobytes(), "strided rerun differs"
        for i, p in enumerate(parts):
            Ai, Bi, Ci = A[i * sa:(i + 1) * sa], B[i * sb:(i + 1) * sb], C[i * sc:(i + 1) * sc]
            gold = oracle.cgemm(ta, tb, m, n, k, alpha, Ai, lda, Bi, ldb, beta, Ci, ldc)
            _check_member(outs[0][i * sc:(i + 1) * sc], gold, Ci, p[6], p[7], alpha, beta, m, n, k, ldc, fills[i],
                          "strided %s member %d %s" % (ops, i, fills[i]))


def test_special_values_grouped(gpu):
    """A grouped batch of 18 members with special fills and all 9 ops: packed members
    (k = 333, correction off; k = 1100, on; a K-split member) and members on the one-launch
    kernel."""
    import torch

    import cgemm3m

    shapes = [(300, 200, 333), (96, 80, 1100), (40, 36, 200), (256, 256, 2048)]
    probs, host = [], []
    for i in range(18):
        ta, tb = OPS[i % 9]
        m, n, k = shapes[i % len(shapes)]
        fills = BATCH_FILLS[i % len(BATCH_FILLS)]
        alpha, beta = (complex(1, 0), complex(0, 1)) if i % 2 else (complex(0.7, -0.3), complex(0.5, -0.25))
        A, lda, B, ldb, C, ldc, opA, opB = _problem(ta, tb, m, n, k, *fills, 9700 + 13 * i, pad=(1, 0, 2))
        dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
        probs.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=alpha, A=ptr(dA), lda=lda, B=ptr(dB), ldb=ldb,
                          beta=beta, C=ptr(dC), ldc=ldc))
        host.append(dict(A=A, B=B, C=C, opA=opA, opB=opB, dA=dA, dB=dB, dC=dC, fills=fills))
    print("grouped special", [(p["ta"] + p["tb"], p["m"], p["n"], p["k"]) for p in probs])
    sys.stdout.flush()
    cgemm3m.grouped(gpu, probs)
    torch.cuda.synchronize()
    first = [h["dC"].cpu().numpy().copy() for h in host]
    for i, (p, h) in enumerate(zip(probs, host)):
        gold = oracle.cgemm(p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], h["A"], p["lda"], h["B"], p["ldb"],
                            p["beta"], h["C"], p["ldc"])
        _check_member(first[i], gold, h["C"], h["opA"], h["opB"], p["alpha"], p["beta"], p["m"], p["n"], p["k"],
                      p["ldc"], h["fills"], "grouped member %d %s%s %s" % (i, p["ta"], p["tb"], h["fills"]))
    for h in host:
        h["dC"].copy_(torch.from_numpy(h["C"]).cuda())
    cgemm3m.grouped(gpu, probs)
    torch.cuda.synchronize()
    for i, h in enumerate(host):
        assert h["dC"].cpu().numpy().tobytes() == first[i].tobytes(), "grouped rerun differs in member %d" % i


@pytest.mark.parametrize("m,n,k", [(32, 32, 1), (64, 40, 40), (300, 200, 333), (200, 96, 1100)])
def test_plane_overflow_comes_back_finite(gpu, m, n, k):
    """A = (3e38, 3e38) times B = (1, 0): the reference BLAS gives (3e38, 3e38); the plane
    Ar + Ai overflows, and every path must still return the finite value."""
    import cgemm3m

    rng = np.random.default_rng(m + n + k)
    opA = ((rng.uniform(-1, 1, (m, k)) + 1j * rng.uniform(-1, 1, (m, k))) * 1e-3).astype(np.complex64)
    opB = ((rng.uniform(-1, 1, (k, n)) + 1j * rng.uniform(-1, 1, (k, n))) * 1e-3).astype(np.complex64)
    opA[:, 0] = complex(BIG, BIG)
    opB[0, :] = 1.0
    for ops in ("NN", "CT", "TN"):
        ta, tb = ops
        lda, ldb = (m if ta == "N" else k), (k if tb == "N" else n)
        A = _store(opA, ta, *((m, k) if ta == "N" else (k, m)), lda)
        B = _store(opB, tb, *((k, n) if tb == "N" else (n, k)), ldb)
        C = np.zeros(m * n, np.complex64)
        for pin in PATHS:
            cfg = cgemm3m.config(**pin) if pin else None
            out, plan = _run(gpu, ta, tb, m, n, k, 1.0, A, lda, B, ldb, 0.0, C, m, cfg)
            gold = oracle.cgemm(ta, tb, m, n, k, 1.0, A, lda, B, ldb, 0.0, C, m)
            assert np.isfinite(gold.view(np.float32)).all()
            fin = np.isfinite(out.view(np.float32)).all()
            assert fin, (ops, m, n, k, pin, plan)
            ok, stats = oracle.verify(out, gold, C, m, n, m)
            assert ok, (ops, pin, plan, stats)


def _segment_problem(size=2048, r=1.0):
    """SegmentCancellation data (cgemm3m_test.cpp:212-240): C(0,0) is exactly r."""
    a = np.zeros(size * size, np.complex64)
    b = np.zeros(size * size, np.complex64)
    for l in range(40 * 33):
        a[l * size] = 1
        b[l] = 4096.1 if l < 20 * 33 else -4096.1
    a[(size - 1) * size] = 1
    b[size - 1] = r
    expected = np.zeros(size * size, np.complex64)
    expected[0] = r
    return a, b, expected


def test_strided_member_segment_cancellation(gpu):
    """A strided batch of three 2048^3 members whose middle one is SegmentCancellation: that
    member returns r within 0.01 and passes the verifier against the exact answer; the
    others are within the FP32 tolerance of the FP64 product; reruns are byte-identical."""
    import torch

    import cgemm3m

    size = 2048
    s = size * size
    a, b, expected = _segment_problem(size, 1.0)
    A = oracle.fill(size, 3 * size, size, 9301)
    B = oracle.fill(size, 3 * size, size, 9302)
    A[s:2 * s], B[s:2 * s] = a, b
    C = np.zeros(3 * s, np.complex64)
    dA, dB = to_dev(A), to_dev(B)
    outs = []
    for _ in range(2):
        dC = to_dev(C)
        cgemm3m.strided_batched(gpu, "N", "N", size, size, size, 1.0, ptr(dA), size, s, ptr(dB), size, s, 0.0,
                                ptr(dC), size, s, 3)
        torch.cuda.synchronize()
        outs.append(dC.cpu().numpy())
    assert outs[0].tobytes() == outs[1].tobytes(), "strided rerun differs"
    mid = outs[0][s:2 * s]
    print("strided segment member C(0,0) =", mid[0])
    assert abs(float(mid[0].real) - 1.0) <= 0.01 and abs(float(mid[0].imag)) <= 0.01, mid[0]
    ok, stats = oracle.verify(mid, expected, np.zeros(s, np.complex64), size, size, size)
    assert ok, stats
    for i in (0, 2):
        dCi = to_dev(outs[0][i * s:(i + 1) * s])
        res = cgemm3m.residual(gpu, "N", "N", size, size, size, 1.0, ctypes_ptr(dA, i * s), size,
                               ctypes_ptr(dB, i * s), size, 0.0, None, ptr(dCi), size)
        assert res["normwise"] < 1e-5, (i, res)  # the FP32 tolerance (3M at 2048^3: about 3e-6)


def ctypes_ptr(t, offset):
    import ctypes

    return ctypes.c_void_p(t.data_ptr() + 8 * offset)


def test_grouped_member_segment_cancellation(gpu):
    """A grouped batch holding SegmentCancellation (2048^3, packed), a 300 x 200 x 1100 member
    (C, N) and a 40 x 36 x 1500 member on the one-launch kernel, all with the correction on:
    the segment member returns r within 0.01, the others pass the oracle; a rerun and two
    hipGraph replays are byte-identical to the first call."""
    import torch

    import cgemm3m

    size = 2048
    a, b, expected = _segment_problem(size, 1.0)
    A1, B1, C1 = oracle.fill(1100, 300, 1100, 9401), oracle.fill(1100, 200, 1100, 9402), oracle.fill(300, 200, 300, 9403)
    A2, B2, C2 = oracle.fill(40, 1500, 40, 9404), oracle.fill(36, 1500, 36, 9405), oracle.fill(40, 36, 40, 9406)
    dev = [to_dev(x) for x in (a, b, A1, B1, A2, B2)]
    C0 = np.zeros(size * size, np.complex64)
    dCs = [to_dev(C0), to_dev(C1), to_dev(C2)]
    probs = [dict(ta="N", tb="N", m=size, n=size, k=size, alpha=1.0, A=ptr(dev[0]), lda=size, B=ptr(dev[1]),
                  ldb=size, beta=0.0, C=ptr(dCs[0]), ldc=size),
             dict(ta="C", tb="N", m=300, n=200, k=1100, alpha=complex(0.7, -0.3), A=ptr(dev[2]), lda=1100,
                  B=ptr(dev[3]), ldb=1100, beta=complex(0.2, 0.9), C=ptr(dCs[1]), ldc=300),
             dict(ta="N", tb="T", m=40, n=36, k=1500, alpha=complex(1, 0.5), A=ptr(dev[4]), lda=40, B=ptr(dev[5]),
                  ldb=36, beta=complex(0.5, -0.25), C=ptr(dCs[2]), ldc=40)]
    inits = [C0, C1, C2]

    def reset():
        for d, c in zip(dCs, inits):
            d.copy_(torch.from_numpy(c).cuda())

    cgemm3m.grouped(gpu, probs)
    torch.cuda.synchronize()
    first = [d.cpu().numpy().copy() for d in dCs]
    seg = first[0]
    print("grouped segment member C(0,0) =", seg[0])
    assert abs(float(seg[0].real) - 1.0) <= 0.01 and abs(float(seg[0].imag)) <= 0.01, seg[0]
    ok, stats = oracle.verify(seg, expected, C0, size, size, size)
    assert ok, stats
    g1 = oracle.cgemm("C", "N", 300, 200, 1100, complex(0.7, -0.3), A1, 1100, B1, 1100, complex(0.2, 0.9), C1, 300)
    g2 = oracle.cgemm("N", "T", 40, 36, 1500, complex(1, 0.5), A2, 40, B2, 36, complex(0.5, -0.25), C2, 40)
    for out, gold, c, m, n in ((first[1], g1, C1, 300, 200), (first[2], g2, C2, 40, 36)):
        ok, worst = oracle.reltol_check(region(out, m, n, m), region(gold, m, n, m))
        verdict, stats = oracle.verify(out, gold, c, m, n, m)
        assert ok and verdict, (m, n, worst, stats)
    reset()
    cgemm3m.grouped(gpu, probs)
    torch.cuda.synchronize()
    for d, f in zip(dCs, first):
        assert d.cpu().numpy().tobytes() == f.tobytes(), "grouped rerun differs"
    s = torch.cuda.Stream()
    gpu.set_stream(s)
    try:
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            cgemm3m.grouped(gpu, probs)
        for _ in range(2):
            reset()
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            for d, f in zip(dCs, first):
                assert d.cpu().numpy().tobytes() == f.tobytes(), "graph replay differs"
    finally:
        gpu.set_stream(None)


def test_fp16_block_with_inf_keeps_other_rows(gpu):
    """FP16 block-scaled products: one Inf in op(A) inside a 128 x 128 block whose other rows
    hold values of about 1e6. The block's scale comes from its finite values, so the rows
    without the Inf stay finite and within FP16's normwise error; the Inf row is non-finite
    where the reference BLAS's is."""
    import cgemm3m

    m, n, k = 256, 192, 256
    rng = np.random.default_rng(77)
    opA = ((rng.uniform(-1, 1, (m, k)) + 1j * rng.uniform(-1, 1, (m, k))) * 1e6).astype(np.complex64)
    opB = (rng.uniform(-1, 1, (k, n)) + 1j * rng.uniform(-1, 1, (k, n))).astype(np.complex64)
    opA[5, 17] = complex(np.inf, 0.0)
    A = _store(opA, "N", m, k, m)
    B = _store(opB, "N", k, n, k)
    C = np.zeros(m * n, np.complex64)
    out, plan = _run(gpu, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, cgemm3m.config("fp16"))
    gold = oracle.cgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    assert _nonfinite_agree(out, gold, m, n, m) == 0, plan
    o, g = region(out, m, n, m), region(gold, m, n, m)
    rows = np.arange(m) != 5
    assert np.isfinite(o[rows].view(np.float32)).all()
    err = np.abs(o[rows].astype(np.complex128) - g[rows]).max() / np.abs(g[rows]).max()
    print("fp16 normwise error without the Inf row", err, plan)
    assert err < 2e-3, err
