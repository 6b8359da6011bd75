"""CPU tests that pin the oracle (oracle/oracle.c) before anything is checked against it.

Each check is chosen so that a plausible mistake in the restatement fails it: a dropped or
sign-flipped term (brute force, integer-exact inputs), a wrong conj/transpose branch
(brute force over all 9 op pairs, conj identity), the wrong loop order or beta placement
(closed-form rounding probes per Netlib branch), and the golden wrapper's quick returns
(reference KATs ZeroProductAllowsNullInputs / EmptyOutputPreservesSentinel).
"""
import itertools
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle

OPS = ["N", "T", "C"]


def _rand_case(ta, tb, m, n, k, seed, pad=1):
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ar + pad, br + pad, m + pad
    A = oracle.fill(ar, ac, lda, seed)
    B = oracle.fill(br, bc, ldb, seed + 1)
    C = oracle.fill(m, n, ldc, seed + 2)
    return A, lda, B, ldb, C, ldc


@pytest.mark.parametrize("ta,tb", list(itertools.product(OPS, OPS)))
def test_bruteforce_all_ops(ta, tb):
    """cgemm within a forward-error bound of the exact rational result; zgemm within 1e-15."""
    m, n, k = 3, 4, 5
    A, lda, B, ldb, C, ldc = _rand_case(ta, tb, m, n, k, seed=11)
    alpha, beta = complex(0.7, -0.3), complex(0.2, 0.9)
    exact = oracle.bruteforce(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    got = oracle.cgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    z = oracle.zgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    u = 2.0 ** -24
    for idx, (er, ei) in enumerate(exact):
        j, i = divmod(idx, m)
        g = got[j * ldc + i]
        # |terms| bound: (|alpha| sum|a||b| + |beta||c|) * 2, times (k + 4) roundings
        bound = (k + 6) * u * 4 * (1 + abs(alpha) * k * 2 + abs(beta) * 2)
        assert abs(float(g.real) - float(er)) <= bound
        assert abs(float(g.imag) - float(ei)) <= bound
        zz = z[j * m + i]
        assert abs(zz.real - float(er)) <= 1e-14 * (1 + abs(float(er)))
        assert abs(zz.imag - float(ei)) <= 1e-14 * (1 + abs(float(ei)))


@pytest.mark.parametrize("ta,tb", list(itertools.product(OPS, OPS)))
def test_integer_inputs_exact(ta, tb):
    """Small-integer inputs keep every partial sum exact in FP32: oracle == exact, bitwise."""
    m, n, k = 5, 6, 7
    rng = np.random.default_rng(3)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    A = (rng.integers(-8, 9, ar * ac) + 1j * rng.integers(-8, 9, ar * ac)).astype(np.complex64)
    B = (rng.integers(-8, 9, br * bc) + 1j * rng.integers(-8, 9, br * bc)).astype(np.complex64)
    C = (rng.integers(-8, 9, m * n) + 1j * rng.integers(-8, 9, m * n)).astype(np.complex64)
    alpha, beta = complex(2, -1), complex(-1, 3)
    exact = oracle.bruteforce(ta, tb, m, n, k, alpha, A, ar, B, br, beta, C, m)
    got = oracle.cgemm(ta, tb, m, n, k, alpha, A, ar, B, br, beta, C, m)
    ref = np.array([complex(float(r), float(i)) for r, i in exact], dtype=np.complex64)
    assert got.tobytes() == ref.tobytes()


def test_netlib_notrans_order_and_beta_first():
    """op(A)=N: C(:,j) is scaled by beta first, then terms are added in l order, one at a time.

    With C=1, beta=1 and two terms 2^-24, sequential addition gives (1+2^-24)+2^-24 = 1
    (each tie rounds to even). The dot-product branch (op(A)=T) sums the terms first:
    2^-24+2^-24 = 2^-23, then adds beta*C: 1+2^-23. The oracle must give exactly these.
    """
    t = np.float32(2.0 ** -24)
    A = np.array([1, 1], dtype=np.complex64)          # 1x2 (N) or 2x1 (T)
    B = np.array([t, t], dtype=np.complex64)          # 2x1
    C = np.array([1], dtype=np.complex64)
    nn = oracle.cgemm("N", "N", 1, 1, 2, 1, A, 1, B, 2, 1, C, 1)
    tn = oracle.cgemm("T", "N", 1, 1, 2, 1, A, 2, B, 2, 1, C, 1)
    assert nn[0] == np.complex64(1.0)
    assert tn[0] == np.complex64(1.0 + 2.0 ** -23)
    # left-to-right (not right-to-left or pairwise) inside one column: 1 + 2^-24 + 2^-24 = 1
    A3 = np.array([1, 1, 1], dtype=np.complex64)
    B3 = np.array([1, t, t], dtype=np.complex64)
    nn3 = oracle.cgemm("N", "N", 1, 1, 3, 1, A3, 1, B3, 3, 0, np.zeros(1, np.complex64), 1)
    tn3 = oracle.cgemm("T", "N", 1, 1, 3, 1, A3, 3, B3, 3, 0, np.zeros(1, np.complex64), 1)
    assert nn3[0] == np.complex64(1.0) and tn3[0] == np.complex64(1.0)
    B3r = np.array([t, t, 1], dtype=np.complex64)
    nn3r = oracle.cgemm("N", "N", 1, 1, 3, 1, A3, 1, B3r, 3, 0, np.zeros(1, np.complex64), 1)
    assert nn3r[0] == np.complex64(1.0 + 2.0 ** -23)


def test_netlib_temp_is_alpha_times_b():
    """op(A)=N branch forms temp = alpha*B(l,j) (rounded) before multiplying A(:,l).

    With alpha = 1+2988u, b = 1+2590u, a = 1+2227u (u = 2^-23) the two association orders
    round differently; the oracle must match fl(fl(alpha*b)*a), computed here with numpy
    float32 scalars.
    """
    u = 2.0 ** -23
    al = np.float32(1 + 2988 * u)
    b = np.float32(1 + 2590 * u)
    a = np.float32(1 + 2227 * u)
    netlib = np.float32(np.float32(al * b) * a)
    other = np.float32(al * np.float32(b * a))
    assert netlib != other  # the probe distinguishes the two orders
    out = oracle.cgemm("N", "N", 1, 1, 1, complex(al, 0), np.array([a], np.complex64), 1,
                       np.array([b], np.complex64), 1, 0, np.zeros(1, np.complex64), 1)
    assert out[0].real == netlib


def test_conj_identity_bitwise():
    """For real alpha and beta = 0: cgemm(C,C) == conj(cgemm(T,T)) bit for bit (negation is exact)."""
    m, n, k = 7, 5, 9
    A, lda, B, ldb, C, ldc = _rand_case("C", "C", m, n, k, seed=5)
    cc = oracle.cgemm("C", "C", m, n, k, 0.75, A, lda, B, ldb, 0, C, ldc)
    tt = oracle.cgemm("T", "T", m, n, k, 0.75, A, lda, B, ldb, 0, C, ldc)
    r_cc = np.array([cc[j * ldc + i] for j in range(n) for i in range(m)])
    r_tt = np.array([tt[j * ldc + i] for j in range(n) for i in range(m)])
    assert r_cc.tobytes() == np.conj(r_tt).astype(np.complex64).tobytes()


def test_parallel_equals_serial():
    for ta, tb in [("N", "N"), ("C", "T"), ("T", "C")]:
        m, n, k = 33, 29, 47
        A, lda, B, ldb, C, ldc = _rand_case(ta, tb, m, n, k, seed=9)
        s = oracle.cgemm(ta, tb, m, n, k, (1, 0.5), A, lda, B, ldb, (0.5, -0.25), C, ldc, threads=1)
        p = oracle.cgemm(ta, tb, m, n, k, (1, 0.5), A, lda, B, ldb, (0.5, -0.25), C, ldc, threads=8)
        assert s.tobytes() == p.tobytes()


def test_zero_product_allows_null_inputs():
    """Reference KAT ZeroProductAllowsNullInputs (cgemm3m_test.cpp:189-210) on the golden."""
    for zero_k in (True, False):
        k = 0 if zero_k else 7
        alpha = 1.0 if zero_k else 0.0
        for beta in (0, 1, complex(0.5, -0.25)):
            inp = np.float32("nan") if beta == 0 else np.float32(7)
            init = np.full(10, complex(inp, -3), dtype=np.complex64)
            out = oracle.cgemm("N", "N", 3, 2, k, alpha, None, 3, None, 7, beta, init, 5)
            for j in range(2):
                for i in range(5):
                    v, v0 = out[j * 5 + i], init[j * 5 + i]
                    if i >= 3:
                        assert v.tobytes() == v0.tobytes()
                    elif beta == 0:
                        assert v == 0
                    elif beta == 1:
                        assert v.tobytes() == v0.tobytes()
                    else:
                        b = np.complex64(beta)
                        er = np.float32(np.float32(b.real * v0.real) - np.float32(b.imag * v0.imag))
                        ei = np.float32(np.float32(b.real * v0.imag) + np.float32(b.imag * v0.real))
                        assert v.real == er and v.imag == ei


def test_empty_output_preserves_sentinel():
    """Reference KAT EmptyOutputPreservesSentinel (cgemm3m_test.cpp:175-187)."""
    sentinel = np.full(12, complex(7, -3), dtype=np.complex64)
    for m, n in ((0, 3), (3, 0)):
        out = oracle.cgemm("N", "N", m, n, 2, 1, np.zeros(8, np.complex64), 4, np.zeros(8, np.complex64), 4, 0,
                           sentinel, 4)
        assert out.tobytes() == sentinel.tobytes()


def test_segment_cancellation_shows_sequential_drift():
    """Reference KAT SegmentCancellation data (cgemm3m_test.cpp:212-240), column 0 only.

    The exact C(0,0) is r; Netlib's sequential FP32 order drifts by 15/1024 because the
    rounding on the way up (660 x 4096.1) is not undone on the way down. That drift,
    reproduced here with a numpy float32 scalar loop, is the oracle's known answer; the
    reference (and the GPU tests) check the product against the exact r instead.
    """
    size = 2048
    a = np.zeros(size * size, np.complex64)
    b = np.zeros(size, np.complex64)
    for l in range(40 * 33):
        a[l * size] = 1
        b[l] = 4096.1 if l < 20 * 33 else -4096.1
    a[(size - 1) * size] = 1
    for r in (0.0, 1.0):
        b[size - 1] = r
        out = oracle.cgemm("N", "N", size, 1, size, 1, a, size, b, size, 0, np.zeros(size, np.complex64), size)
        s = np.float32(0)
        for l in range(size):
            s = np.float32(s + np.float32(b[l].real) * np.float32(a[l * size].real))
        assert out[0].real == s
        assert abs(float(s) - (r + 15 / 1024)) == 0.0
        assert out[1] == 0


def test_verifier_kats():
    """The reference's verifier KATs (cgemm3m_test.cpp:93-164) on oracle_verify_component."""
    V = oracle.verify_component
    z99 = [0.0] * 99
    assert V(z99, z99)[0]
    assert V(z99 + [0.001], z99 + [0.0])[0]
    assert not V(z99 + [0.001, 0.001], z99 + [0.0, 0.0])[0]
    assert V(z99 + [0.001, 0.001] + z99, z99 + [0.0, 0.0] + z99)[0]
    assert not V(z99 + [0.02], z99 + [0.0])[0]
    assert not V([0.001], [0.0])[0]
    ok, st = V([0.001, 8192.02], [0.0, 8192.0])
    assert not ok and st[6] == 8192.0
    assert V([1.0 / 8192.0], [0.0])[0]
    assert not V([1.0 / 8192.0, 2.0 / 8192.0], [0.0, 0.0])[0]
    ok, st = V([8192.5], [8192.0])
    assert st[1] == 1 and not ok
    ok, st = V([8192.03125], [8192.0])
    assert st[1] == st[0] and ok
    ok, st = V([8192.03125, 8192.0625], [8192.0, 8192.0])
    assert st[1] == 2 and not ok
    inf, nan = float("inf"), float("nan")
    assert V([nan, inf, -inf], [nan, inf, -inf])[0]
    for a, g in ((nan, inf), (inf, nan), (-inf, nan), (nan, -inf)):
        assert V([a], [g])[0]
    for a, g in ((inf, -inf), (-inf, inf)):
        assert not V([a], [g])[0]
    assert not V([0.0], [inf])[0]


def test_prng_streams_match():
    a = np.zeros(1000, dtype=np.float32)
    oracle.lib().oracle_fill(1234, 77, 1000, oracle._fp(a))
    b = oracle.uniform(1234, 77, 1000)
    assert a.tobytes() == b.tobytes()
    assert b.min() >= -1.0 and b.max() < 1.0
    assert abs(float(b.mean())) < 0.1


def test_reltol_check_semantics():
    g = np.array([1 + 1j, 100 + 0j], dtype=np.complex64)
    assert oracle.reltol_check(g, g)[0]
    x = g.copy()
    x[0] += np.complex64(0.5e-3)  # bound for component 0: 1e-5*1 + 1e-5*100 = 1.01e-3
    assert oracle.reltol_check(x, g)[0]
    x[0] = g[0] + np.complex64(2e-3)
    assert not oracle.reltol_check(x, g)[0]


def test_bruteforce_is_exact_rational():
    out = oracle.bruteforce("N", "N", 1, 1, 2, 1, np.array([0.1, 0.2], np.complex64), 1,
                            np.array([3, 4], np.complex64), 2, 0, np.zeros(1, np.complex64), 1)
    want = Fraction(float(np.float32(0.1))) * 3 + Fraction(float(np.float32(0.2))) * 4
    assert out[0][0] == want and out[0][1] == 0


def _golden():
    import json

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    data = np.load(os.path.join(here, "golden.npz"))  # plain arrays, no pickle
    with open(os.path.join(here, "golden.json")) as f:
        meta = json.load(f)
    return data, meta


def test_golden_fixtures_reproduce():
    """The committed fixtures (tests/golden/make_golden.py): the oracle reproduces every stored
    FP32 result bitwise and FP64 result exactly; the FP64 results agree with numpy's complex128
    product of the same operands (an independent library routine) to 1e-12 relative; and the
    FP32 results sit within the FP32 error bound of the FP64 ones."""
    data, meta = _golden()
    assert len(meta) == 13
    for i, c in enumerate(meta):
        ta, tb, m, n, k = c["ta"], c["tb"], c["m"], c["n"], c["k"]
        lda, ldb, ldc = c["lda"], c["ldb"], c["ldc"]
        alpha, beta = complex(*c["alpha"]), complex(*c["beta"])
        A, B, C = data["%d_A" % i], data["%d_B" % i], data["%d_C" % i]
        out32 = oracle.cgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
        assert out32.tobytes() == data["%d_out32" % i].tobytes(), i
        out64 = oracle.zgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
        assert np.array_equal(out64, data["%d_out64" % i]), i
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        opA = oracle.op_matrix(A, ar, ac, lda, ta)
        opB = oracle.op_matrix(B, br, bc, ldb, tb)
        al = complex(np.float32(alpha.real), np.float32(alpha.imag))
        be = complex(np.float32(beta.real), np.float32(beta.imag))
        ref = al * (opA @ opB)
        if be != 0:
            ref = ref + be * C[: ldc * n].reshape(n, ldc).T[:m].astype(np.complex128)
        got64 = out64.reshape(n, m).T
        assert np.allclose(got64, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), i
        got32 = out32[: ldc * n].reshape(n, ldc).T[:m].astype(np.complex128)
        bound = 4 * (k + 2) * 2.0 ** -24 * (np.abs(al) * (np.abs(opA) @ np.abs(opB)) +
                                            (np.abs(be) * np.abs(C[: ldc * n].reshape(n, ldc).T[:m]) if be != 0 else 0))
        assert (np.abs(got32 - ref) <= bound + 1e-30).all(), i
