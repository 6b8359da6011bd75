"""A functional case list in the categories of the reference's accuracy suite, regenerated here
by a seeded generator of this project (no case is copied from the reference's CSV).

The reference's suite (ops-blas/test/gemm/cgemm3m/arch35/cgemm3m_test.csv, 1000 functional
rows) groups its cases as: squares 1..2048 (SQ), non-square shapes (MN), a large random sweep
over shapes, ops, scalars and padded leading dimensions (EX), scalar sets on 32^3 (AB), padded
leading dimensions (LD), complex scalar values (CV); special values (FL) and the edge contract
(ED) have their own tests. Every case here is judged the way the reference judges its rows:
the verifier against the Netlib-order oracle (padding rows included) and a byte-identical
second run.

A case is a dict: cat, ta, tb, m, n, k, alpha, beta, lda, ldb, ldc, seed.
"""
import itertools
import math
import random

OPS = ["".join(p) for p in itertools.product("NTC", "NTC")]
# (alpha, beta): identity, accumulate, general complex, scale only, negated product
SCALARS = [(1, 0), (1, 1), (complex(0.5, 0.5), complex(0.5, 0.25)), (0, 1), (-1, 1)]
# complex scalar values: pure imaginary, general, negative real
CV_SCALARS = [(1, 1j), (1j, 0), (complex(0.7, -0.3), complex(0.2, 0.9)), (-1j, complex(0, -1)),
              (complex(-0.25, 2.0), -0.5)]


def _ld(ta, rows, cols, pad):
    """Leading dimension of the stored matrix whose op() is rows x cols."""
    return max(1, (rows if ta == "N" else cols) + pad)


def _case(cat, ops, m, n, k, scal, pad, seed):
    ta, tb = ops[0], ops[1]
    alpha, beta = scal
    return dict(cat=cat, ta=ta, tb=tb, m=m, n=n, k=k, alpha=complex(alpha), beta=complex(beta),
                lda=_ld(ta, m, k, pad[0]), ldb=_ld(tb, k, n, pad[1]), ldc=max(1, m + pad[2]), seed=seed)


def _logint(rng, lo, hi):
    """Integer in [lo, hi], log-uniform."""
    return int(min(hi, max(lo, round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))))


def suite_cases(seed=20261204):
    rng = random.Random(seed)
    cases = []
    s = seed * 10
    # squares: sizes around powers of two and odd sizes, each op pair on a rotating subset
    sq_sizes = [1, 2, 3, 4, 5, 7, 8, 9, 15, 16, 17, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 200, 255, 256, 257,
                384, 500, 511, 512, 513, 640, 768, 1000, 1023, 1024, 1025, 1536, 2048]
    for i, sz in enumerate(sq_sizes):
        reps = 3 if sz <= 1024 else 1
        for r in range(reps):
            s += 1
            cases.append(_case("square", OPS[(3 * i + r) % 9], sz, sz, sz, SCALARS[(i + r) % len(SCALARS)],
                               (0, 0, 0), s))
    # non-square shapes: m, n, k independently small, odd or around tile edges
    mn_vals = [1, 3, 10, 16, 31, 33, 65, 96, 130, 257, 500, 1001]
    for i in range(81):
        s += 1
        m, n, k = rng.choice(mn_vals), rng.choice(mn_vals), rng.choice(mn_vals + [1024])
        cases.append(_case("rect", OPS[i % 9], m, n, k, SCALARS[i % len(SCALARS)], (0, 0, 0), s))
    # random sweep: log-uniform sizes to 1024, any op pair, scalar set and padding
    for i in range(663):
        s += 1
        m, n, k = _logint(rng, 1, 1024), _logint(rng, 1, 1024), _logint(rng, 1, 1024)
        pad = tuple(rng.choice([0, 0, 1, 3, 8, 13]) for _ in range(3))
        cases.append(_case("sweep", rng.choice(OPS), m, n, k, rng.choice(SCALARS), pad, s))
    # scalar sets on 32^3, every op pair
    for i, (ops, scal) in enumerate(itertools.product(OPS, SCALARS[:4])):
        s += 1
        cases.append(_case("scalars", ops, 32, 32, 32, scal, (0, 0, 0), s))
    # padded leading dimensions
    for i, sz in enumerate([16, 32, 64, 128]):
        for j, ops in enumerate(["NN", "TC", "CT", "NT"]):
            s += 1
            cases.append(_case("ld", ops, sz, sz, sz, SCALARS[0], (4, 8, 4 + 3 * j), s))
    # complex scalar values
    for i, sz in enumerate([10, 33, 100, 400]):
        for j in range(8):
            s += 1
            cases.append(_case("complex_scalars", OPS[(i + j) % 9], sz, sz + j, sz, CV_SCALARS[j % len(CV_SCALARS)],
                               (0, 1, 2), s))
    return cases


def chunks(cases, size=50):
    return [cases[i:i + size] for i in range(0, len(cases), size)]
