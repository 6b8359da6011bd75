"""Operands beyond 2^31 elements: 64-bit addressing on every kernel of the path.

op(A) is 65600 x 32800 (2.15e9 complex elements, 17.2 GB, more than 2^31 elements and 2^34
bytes), B 32800 x 100, C 65600 x 100 with ldc = 65603. The same device buffer serves as A
stored 65600 x 32800 (op N, lda = 65600) and as A stored 32800 x 65600 (op C, lda = 32800).
Each call is checked on sampled output rows, the last row among them (its operand row ends
at the buffer's last element for op N). FP32: against the oracle restricted to those rows (in
the reference BLAS's order every output depends only on its own row of op(A) and column of
op(B), so the sub-problem's outputs are the full problem's) and against FP64. At k = 32800
the oracle's own sequential FP32 sums are off by about 1e-3 in absolute terms, more than the
reference verifier's per-output allowance (1 + |golden|) / 8192 for outputs near zero, so the
bar is the one of the tall-skinny parity test: error against FP64 no worse than twice the
oracle's, and the reltol check against the oracle at the larger of 1e-5 and twice the
oracle's normwise error. 16-bit forms: against FP64 with their normwise bound. The padding
rows of C stay untouched and reruns are byte-identical.
The strided batched entry takes a member whose A lies 2^31 + 64 elements into the buffer, and
the grouped entry takes the large problem as a member next to a small one."""
import sys

import numpy as np
import pytest

import oracle
from helpers import check_against_oracle, ptr

pytestmark = pytest.mark.gpu

M, N, K = 65600, 100, 32800
LDC = M + 3
ROWS = [0, 1, 31337, M - 65, M - 1]
ALPHA, BETA = complex(0.7, -0.3), complex(0.2, 0.9)
BOUND = {"bf16x3": 6e-5, "fp16": 4e-3}


def _say(*a):
    print(*a)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def big(gpu):
    import torch

    import cgemm3m

    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info()
    if free < (96 << 30):
        pytest.skip("needs 96 GB of free device memory, %.1f GB free" % (free / 2 ** 30))
    g = torch.Generator(device="cuda")
    g.manual_seed(20261204)

    def u(count):
        return (torch.rand(2 * count, generator=g, device="cuda") * 2 - 1).view(torch.complex64)

    d = dict(A=u(M * K), B=u(K * N), C=u(LDC * N))
    # a handle of its own: its workspace (up to 26 GB here) is released at the end
    h = cgemm3m.Handle()
    d["h"] = h
    d["Bh"] = d["B"].cpu().numpy()
    yield d
    h.close()
    del d
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _op_rows(d, ta, rows):
    """op(A) restricted to `rows`, as the stored matrix of a sub-problem with the same op
    (column-major, leading dimension len(rows) for op N, K for op C)."""
    A = d["A"]
    if ta == "N":  # stored M x K, lda = M: element (i, l) at l * M + i
        return A.view(K, M)[:, rows].contiguous().cpu().numpy().reshape(-1), len(rows)
    # stored K x M, lda = K: op(A)(i, l) = conj(A(l, i)) at i * K + l
    return A.view(M, K)[rows, :].contiguous().cpu().numpy().reshape(-1), K


def _call(d, ta, cfg=None):
    """One call on a fresh copy of C; returns the output buffer (device)."""
    import cgemm3m

    lda = M if ta == "N" else K
    C = d["C"].clone()
    st = cgemm3m.cgemm3m(d["h"], ta, "N", M, N, K, ALPHA, ptr(d["A"]), lda, ptr(d["B"]), K, BETA, ptr(C), LDC, cfg)
    assert st == 0, st
    return C


def _sampled(C, rows):
    """Rows of the column-major M x N output (ldc = LDC) as a column-major len(rows) x N array."""
    return C.view(N, LDC)[:, rows].contiguous().cpu().numpy().reshape(-1)


def _padding_untouched(C0, C1):
    import torch

    return torch.equal(C0.view(N, LDC)[:, M:], C1.view(N, LDC)[:, M:])


def _fp64_rows(d, ta, rows):
    """alpha * op(A)[rows, :] @ B + beta * C[rows, :] in complex128."""
    a, lda = _op_rows(d, ta, rows)
    r = len(rows)
    if ta == "N":
        opA = a.reshape(K, r).T.astype(np.complex128)
    else:
        opA = np.conj(a.reshape(r, K)).astype(np.complex128)
    B = d["Bh"].reshape(N, K).T.astype(np.complex128)
    C0 = _sampled(d["C"], rows).reshape(N, r).T.astype(np.complex128)
    return ALPHA * (opA @ B) + BETA * C0


def _check_rows(d, ta, C1, label):
    """The FP32 bar on the sampled rows: FP64 error at most twice the oracle's, reltol against
    the oracle at max(1e-5, twice the oracle's normwise error)."""
    a, lda = _op_rows(d, ta, ROWS)
    r = len(ROWS)
    C0 = _sampled(d["C"], ROWS)
    gold = oracle.cgemm(ta, "N", r, N, K, ALPHA, a, lda, d["Bh"], K, BETA, C0, r)
    got = _sampled(C1, ROWS)
    ref = _fp64_rows(d, ta, ROWS)

    def err(x):
        return np.linalg.norm(x.reshape(N, r).T.astype(np.complex128) - ref) / np.linalg.norm(ref)

    ours, netlib = err(got), err(gold)
    tol = max(1e-5, 2 * netlib)
    ok, worst = oracle.reltol_check(got, gold, rtol=tol, atol_frac=tol)
    verdict, stats = oracle.verify(got, gold, C0, r, N, r)
    _say("%s: FP64 normwise ours %.3g oracle %.3g; reltol at %.3g worst %.3g; verifier %s %s" %
         (label, ours, netlib, tol, worst, verdict, stats))
    assert np.isfinite(got).all()
    assert ours <= 2 * netlib, (ours, netlib)
    assert ok, worst


@pytest.mark.parametrize("ta", ["N", "C"])
def test_fp32_operand_beyond_2g_elements(gpu, big, ta):
    d = big
    _say("case large fp32 op %s: %d x %d x %d, op(A) %.1f GB" % (ta, M, N, K, M * K * 8 / 1e9))
    C1 = _call(d, ta)
    plan = d["h"].last_plan(sync=True)
    _say("plan", plan)
    C2 = _call(d, ta)
    import torch

    torch.cuda.synchronize()
    assert torch.equal(C1, C2), "rerun differs"
    assert _padding_untouched(d["C"], C1), "padding rows of C were written"
    _check_rows(d, ta, C1, "large fp32 %s" % ta)


@pytest.mark.parametrize("prec", ["bf16x3", "fp16"])
def test_16bit_operand_beyond_2g_elements(gpu, big, prec):
    import torch

    import cgemm3m

    d = big
    _say("case large %s op N: %d x %d x %d" % (prec, M, N, K))
    cfg = cgemm3m.config(prec)
    C1 = _call(d, "N", cfg)
    _say("plan", d["h"].last_plan(sync=True))
    C2 = _call(d, "N", cfg)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2), "rerun differs"
    assert _padding_untouched(d["C"], C1), "padding rows of C were written"
    ref = _fp64_rows(d, "N", ROWS)
    got = _sampled(C1, ROWS).reshape(N, len(ROWS)).T.astype(np.complex128)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    _say("%s normwise error on the sampled rows: %.3g" % (prec, err))
    assert np.isfinite(got).all()
    assert err <= BOUND[prec], err


def test_strided_member_beyond_2g_elements(gpu, big):
    """Two members of 128 x 100 x 2048 (op N, N) whose A lie 0 and 2^31 + 64 elements into the
    buffer; B shared (stride 0); both members against the oracle."""
    import torch

    import cgemm3m

    d = big
    m, n, k = 128, 100, 2048
    stride_a = (1 << 31) + 64
    assert stride_a + m * k <= M * K
    _say("case large strided: member A at element offsets 0 and %d" % stride_a)
    C = torch.zeros(2 * m * n, dtype=torch.complex64, device="cuda")
    C0 = d["C"][: 2 * m * n].clone()
    C.copy_(C0)
    st = cgemm3m.strided_batched(d["h"], "N", "N", m, n, k, ALPHA, ptr(d["A"]), m, stride_a, ptr(d["B"]), k, 0,
                                 BETA, ptr(C), m, m * n, 2)
    assert st == 0
    torch.cuda.synchronize()
    out = C.cpu().numpy()
    c0 = C0.cpu().numpy()
    for b, off in enumerate((0, stride_a)):
        a = d["A"][off: off + m * k].cpu().numpy()
        gold = oracle.cgemm("N", "N", m, n, k, ALPHA, a, m, d["Bh"], k, BETA, c0[b * m * n:(b + 1) * m * n], m)
        check_against_oracle(out[b * m * n:(b + 1) * m * n], gold, c0[b * m * n:(b + 1) * m * n], m, n, m,
                             "strided member %d" % b)


def test_grouped_member_beyond_2g_elements(gpu, big):
    """A grouped call of the large problem (op N) and a 96 x 80 x 1536 problem (op T, C): the
    large member's output is aclblasCgemm3m's, byte for byte (the grouped kernel issues the
    one-problem kernel's MFMAs in its order, and a batch that fills the GPU takes no K split);
    its sampled rows and the small member against the oracle."""
    import torch

    import cgemm3m

    d = big
    _say("case large grouped: %d x %d x %d + 96 x 80 x 1536" % (M, N, K))
    C1 = d["C"].clone()
    m2, n2, k2 = 96, 80, 1536
    A2 = d["A"][-k2 * m2:].clone()  # stored k2 x m2 (op T), the buffer's last elements
    B2 = d["A"][: n2 * k2].clone()  # stored n2 x k2 (op C)
    C2 = d["C"][: m2 * n2].clone()
    C2init = C2.cpu().numpy()
    probs = [dict(ta="N", tb="N", m=M, n=N, k=K, alpha=ALPHA, A=ptr(d["A"]), lda=M, B=ptr(d["B"]), ldb=K, beta=BETA,
                  C=ptr(C1), ldc=LDC),
             dict(ta="T", tb="C", m=m2, n=n2, k=k2, alpha=ALPHA, A=ptr(A2), lda=k2, B=ptr(B2), ldb=n2, beta=BETA,
                  C=ptr(C2), ldc=m2)]
    cgemm3m.grouped(d["h"], probs)
    torch.cuda.synchronize()
    assert _padding_untouched(d["C"], C1), "padding rows of C were written"
    _check_rows(d, "N", C1, "grouped large member")
    single = _call(d, "N")
    torch.cuda.synchronize()
    assert torch.equal(single, C1), "grouped member differs from the single call"
    gold2 = oracle.cgemm("T", "C", m2, n2, k2, ALPHA, A2.cpu().numpy(), k2, B2.cpu().numpy(), n2, BETA, C2init, m2)
    check_against_oracle(C2.cpu().numpy(), gold2, C2init, m2, n2, m2, "grouped small member")
