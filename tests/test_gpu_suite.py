"""The generated functional suite (tests/suite_cases.py) through aclblasCgemm3m: each case
against the Netlib-order oracle with the reference verifier and the reltol check (ldc padding
untouched), and a byte-identical second run, as the reference's accuracy suite judges its rows
(ops-blas/test/gemm/cgemm3m/arch35/cgemm3m_test.cpp:282-310)."""
import sys

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
