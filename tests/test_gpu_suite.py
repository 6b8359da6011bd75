"""The generated functional suite (tests/suite_cases.py) through aclblasCgemm3m: each case
against the Netlib-order oracle with the reference verifier and the reltol check (ldc padding
untouched), and a byte-identical second run, as the reference's accuracy suite judges its rows
(ops-blas/test/gemm/cgemm3m/arch35/cgemm3m_test.cpp:282-310)."""
import sys

import numpy as np
import pytest

import oracle
from helpers import check_against_oracle, inputs, ptr, run_gpu, to_dev
from suite_cases import chunks, suite_cases

pytestmark = pytest.mark.gpu

CHUNKS = chunks(suite_cases())


@pytest.mark.parametrize("chunk", range(len(CHUNKS)))
def test_generated_suite(gpu, chunk):
    for c in CHUNKS[chunk]:
        ta, tb, m, n, k = c["ta"], c["tb"], c["m"], c["n"], c["k"]
        label = "%s %s%s %dx%dx%d ld=(%d,%d,%d) alpha=%s beta=%s" % (c["cat"], ta, tb, m, n, k, c["lda"], c["ldb"],
                                                                   c["ldc"], c["alpha"], c["beta"])
        print("case", label)
        sys.stdout.flush()
        A, B, C = inputs(ta, tb, m, n, k, c["lda"], c["ldb"], c["ldc"], c["seed"])
        dA, dB = to_dev(A), to_dev(B)
        st, out = run_gpu(gpu, ta, tb, m, n, k, c["alpha"], A, c["lda"], B, c["ldb"], c["beta"], C, c["ldc"], None,
                          dA, dB)
        assert st == 0, (label, st)
        gold = oracle.cgemm(ta, tb, m, n, k, c["alpha"], A, c["lda"], B, c["ldb"], c["beta"], C, c["ldc"])
        check_against_oracle(out, gold, C, m, n, c["ldc"], label)
        st2, out2 = run_gpu(gpu, ta, tb, m, n, k, c["alpha"], A, c["lda"], B, c["ldb"], c["beta"], C, c["ldc"], None,
                            dA, dB)
        same = st2 == 0 and out2.tobytes() == out.tobytes()
        assert same, "%s: second run differs" % label


SWEEP = [c for c in suite_cases() if c["cat"] == "sweep"]
SWEEP_CALLS = chunks(SWEEP, 166)


@pytest.mark.parametrize("call", range(len(SWEEP_CALLS)))
def test_generated_sweep_grouped(gpu, call):
    """The sweep's cases (sizes 1 to 1024, every op pair, scalar sets with alpha = 0 among
    them, padded leading dimensions) as the members of grouped calls of about 166: every
    member against the oracle with the verifier, and a second call byte-identical."""
    import torch

    import cgemm3m

    cases = SWEEP_CALLS[call]
    print("grouped call %d: %d members" % (call, len(cases)))
    sys.stdout.flush()
    probs, host = [], []
    for c in cases:
        A, B, C = inputs(c["ta"], c["tb"], c["m"], c["n"], c["k"], c["lda"], c["ldb"], c["ldc"], c["seed"])
        dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
        probs.append(dict(ta=c["ta"], tb=c["tb"], m=c["m"], n=c["n"], k=c["k"], alpha=c["alpha"], A=ptr(dA),
                          lda=c["lda"], B=ptr(dB), ldb=c["ldb"], beta=c["beta"], C=ptr(dC), ldc=c["ldc"]))
        host.append(dict(A=A, B=B, C=C, dC=dC, keep=(dA, dB)))
    cgemm3m.grouped(gpu, probs)
    torch.cuda.synchronize()
    first = [h["dC"].cpu().numpy() for h in host]
    for c, h, out in zip(cases, host, first):
        gold = oracle.cgemm(c["ta"], c["tb"], c["m"], c["n"], c["k"], c["alpha"], h["A"], c["lda"], h["B"], c["ldb"],
                            c["beta"], h["C"], c["ldc"])
        label = "grouped %s%s %dx%dx%d alpha=%s beta=%s" % (c["ta"], c["tb"], c["m"], c["n"], c["k"], c["alpha"],
                                                           c["beta"])
        check_against_oracle(out, gold, h["C"], c["m"], c["n"], c["ldc"], label)
        h["dC"].copy_(torch.from_numpy(h["C"]).cuda())
    cgemm3m.grouped(gpu, probs)
    torch.cuda.synchronize()
    for i, (h, out) in enumerate(zip(host, first)):
        same = h["dC"].cpu().numpy().tobytes() == out.tobytes()
        assert same, "member %d: second call differs" % i


STRIDED = SWEEP[::16]


def _buffer(count, seed):
    return oracle.uniform(seed, 0, 2 * max(count, 1)).view(np.complex64).copy()


@pytest.mark.parametrize("part", range(2))
def test_generated_sweep_strided(gpu, part):
    """Every 16th sweep case through aclblasCgemm3mStridedBatched with 1 to 4 batches: input
    strides of zero (shared operand) or the stored extent plus a gap, output strides past the
    output's ldc x n span (disjoint outputs). Each batch against the oracle with the verifier; the
    gaps between outputs untouched; a second call byte-identical."""
    import torch

    import cgemm3m

    rng = np.random.default_rng(20261205 + part)
    for idx, c in enumerate(STRIDED[part::2]):
        ta, tb, m, n, k = c["ta"], c["tb"], c["m"], c["n"], c["k"]
        lda, ldb, ldc = c["lda"], c["ldb"], c["ldc"]
        batch = 1 + (idx % 4)
        ext_a = lda * ((k if ta == "N" else m) - 1) + (m if ta == "N" else k)
        ext_b = ldb * ((n if tb == "N" else k) - 1) + (k if tb == "N" else n)
        span_c = ldc * n  # the output with its padding rows
        sa = 0 if rng.random() < 0.25 else ext_a + int(rng.integers(0, 9))
        sb = 0 if rng.random() < 0.25 else ext_b + int(rng.integers(0, 9))
        sc = span_c + int(rng.integers(0, 9))
        label = "strided %s%s %dx%dx%d batch=%d strides=(%d,%d,%d) alpha=%s beta=%s" % (
            ta, tb, m, n, k, batch, sa, sb, sc, c["alpha"], c["beta"])
        print("case", label)
        sys.stdout.flush()
        Ab = _buffer(sa * (batch - 1) + ext_a, c["seed"])
        Bb = _buffer(sb * (batch - 1) + ext_b, c["seed"] + 1)
        Cb = _buffer(sc * (batch - 1) + span_c, c["seed"] + 2)
        dA, dB = to_dev(Ab), to_dev(Bb)
        outs = []
        for _ in range(2):
            dC = to_dev(Cb)
            cgemm3m.strided_batched(gpu, ta, tb, m, n, k, c["alpha"], ptr(dA), lda, sa, ptr(dB), ldb, sb,
                                    c["beta"], ptr(dC), ldc, sc, batch)
            torch.cuda.synchronize()
            outs.append(dC.cpu().numpy())
        assert outs[0].tobytes() == outs[1].tobytes(), "%s: second call differs" % label
        out = outs[0]
        touched = np.zeros(Cb.size, dtype=bool)
        for i in range(batch):
            a, b, c0 = Ab[i * sa: i * sa + ext_a], Bb[i * sb: i * sb + ext_b], Cb[i * sc: i * sc + span_c]
            gold = oracle.cgemm(ta, tb, m, n, k, c["alpha"], a, lda, b, ldb, c["beta"], c0, ldc)
            check_against_oracle(out[i * sc: i * sc + span_c], gold, c0, m, n, ldc, "%s batch %d" % (label, i))
            touched[i * sc: i * sc + span_c] = True
        gaps = out[~touched].tobytes() == Cb[~touched].tobytes()
        assert gaps, "%s: output gaps were written" % label


ALPHA_X = complex(1.0, 0.5)
# normwise error bound against FP64 of each 16-bit form (uniform [-1, 1) inputs, k <= 1024)
FORM_BOUND = {"bf16x3": 6e-5, "bf16x6": 4e-6, "fp16": 4e-3}


@pytest.mark.parametrize("prec", sorted(FORM_BOUND))
def test_generated_sweep_precision_forms(gpu, prec):
    """Every 11th sweep case on a 16-bit form: normwise error against FP64 (the GPU residual)
    within the form's bound, nothing non-finite, a second run byte-identical."""
    import torch

    import cgemm3m

    cfg = cgemm3m.config(prec)
    for c in SWEEP[::11]:
        ta, tb, m, n, k = c["ta"], c["tb"], c["m"], c["n"], c["k"]
        label = "%s %s%s %dx%dx%d alpha=%s beta=%s" % (prec, ta, tb, m, n, k, c["alpha"], c["beta"])
        print("case", label)
        sys.stdout.flush()
        A, B, C = inputs(ta, tb, m, n, k, c["lda"], c["ldb"], c["ldc"], c["seed"])
        dA, dB = to_dev(A), to_dev(B)
        st, out = run_gpu(gpu, ta, tb, m, n, k, c["alpha"], A, c["lda"], B, c["ldb"], c["beta"], C, c["ldc"], cfg,
                          dA, dB)
        assert st == 0, (label, st)
        st2, out2 = run_gpu(gpu, ta, tb, m, n, k, c["alpha"], A, c["lda"], B, c["ldb"], c["beta"], C, c["ldc"], cfg,
                            dA, dB)
        assert st2 == 0 and out2.tobytes() == out.tobytes(), "%s: second run differs" % label
        assert np.isfinite(out).all(), label
        dC0, dOut = to_dev(C), to_dev(out)
        res = cgemm3m.residual(gpu, ta, tb, m, n, k, c["alpha"], ptr(dA), c["lda"], ptr(dB), c["ldb"], c["beta"],
                               ptr(dC0), ptr(dOut), c["ldc"])
        torch.cuda.synchronize()
        assert res["normwise"] <= FORM_BOUND[prec], (label, res)


def test_generated_sweep_four_product_cgemm(gpu):
    """Every 8th sweep case through aclblasCgemm (four real products): against the oracle with
    the verifier, padding rows untouched, a second call byte-identical."""
    import torch

    import cgemm3m

    for c in SWEEP[::8]:
        ta, tb, m, n, k = c["ta"], c["tb"], c["m"], c["n"], c["k"]
        label = "cgemm %s%s %dx%dx%d alpha=%s beta=%s" % (ta, tb, m, n, k, c["alpha"], c["beta"])
        print("case", label)
        sys.stdout.flush()
        A, B, C = inputs(ta, tb, m, n, k, c["lda"], c["ldb"], c["ldc"], c["seed"])
        dA, dB = to_dev(A), to_dev(B)
        outs = []
        for _ in range(2):
            dC = to_dev(C)
            st = cgemm3m.cgemm(gpu, ta, tb, m, n, k, c["alpha"], ptr(dA), c["lda"], ptr(dB), c["ldb"], c["beta"],
                               ptr(dC), c["ldc"], check=False)
            assert st == 0, (label, st)
            torch.cuda.synchronize()
            outs.append(dC.cpu().numpy())
        assert outs[0].tobytes() == outs[1].tobytes(), "%s: second call differs" % label
        gold = oracle.cgemm(ta, tb, m, n, k, c["alpha"], A, c["lda"], B, c["ldb"], c["beta"], C, c["ldc"])
        check_against_oracle(outs[0], gold, C, m, n, c["ldc"], label)


def test_generated_sweep_user_workspace(gpu):
    """Every 10th sweep case on a handle given a 48 MiB user workspace (aclblasSetWorkspace):
    the output is byte-identical to the default handle's, or, where the plan needs more than
    the workspace holds, the call returns EXECUTION_FAILED and leaves C alone."""
    import ctypes

    import torch

    import cgemm3m
    from cgemm3m import _lib

    ws_bytes = 48 << 20
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    h = cgemm3m.Handle()
    try:
        assert h.L.aclblasSetWorkspace(h.h, ctypes.c_void_p(ws.data_ptr()), ctypes.c_size_t(ws_bytes)) == 0
        served = refused = 0
        for c in SWEEP[::10]:
            ta, tb, m, n, k = c["ta"], c["tb"], c["m"], c["n"], c["k"]
            A, B, C = inputs(ta, tb, m, n, k, c["lda"], c["ldb"], c["ldc"], c["seed"])
            dA, dB = to_dev(A), to_dev(B)
            st, ref = run_gpu(gpu, ta, tb, m, n, k, c["alpha"], A, c["lda"], B, c["ldb"], c["beta"], C, c["ldc"],
                              None, dA, dB)
            assert st == 0
            need = cgemm3m.workspace_size(h, ta, tb, m, n, k, c["alpha"], c["beta"])
            st, out = run_gpu(h, ta, tb, m, n, k, c["alpha"], A, c["lda"], B, c["ldb"], c["beta"], C, c["ldc"],
                              None, dA, dB)
            if need > ws_bytes:
                assert st == _lib.EXECUTION_FAILED, (c, need, st)
                assert out.tobytes() == C.tobytes(), "a refused call wrote C"
                refused += 1
                continue
            assert st == 0, (c, need, st)
            served += 1
            assert out.tobytes() == ref.tobytes(), "user-workspace output differs: %s" % (c,)
        # a plan beyond the workspace: 2048^3 packs 100 MB of planes
        A, B, C = inputs("N", "N", 2048, 2048, 2048, 2048, 2048, 2048, 20261800)
        need = cgemm3m.workspace_size(h, "N", "N", 2048, 2048, 2048, ALPHA_X, 0.0)
        assert need > ws_bytes
        st, out = run_gpu(h, "N", "N", 2048, 2048, 2048, ALPHA_X, A, 2048, B, 2048, 0.0, C, 2048)
        assert st == _lib.EXECUTION_FAILED, st
        assert out.tobytes() == C.tobytes(), "a refused call wrote C"
        refused += 1
        print("user workspace: %d served, %d refused" % (served, refused))
        assert served >= 50
    finally:
        h.close()
