"""The generated functional suite (tests/suite_cases.py) through aclblasCgemm3m: each case
against the Netlib-order oracle with the reference verifier and the reltol check (ldc padding
untouched), and a byte-identical second run, as the reference's accuracy suite judges its rows
(ops-blas/test/gemm/cgemm3m/arch35/cgemm3m_test.cpp:282-310)."""
import sys

import pytest

import oracle
from helpers import check_against_oracle, inputs, run_gpu, to_dev
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
