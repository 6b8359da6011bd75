"""Seeded random shapes through the planner's default choices, where the rules for classes the
tuned table does not list decide the plan: the FP32 packed path's K split for small outputs with
long k, the one-launch kernel for classes between two one-launch k classes, and the 16-bit forms'
K splits for small outputs. Every case is checked against the oracle (FP32: the reltol check and
the reference verifier, untouched ldc padding) or against FP64 on the GPU (16-bit forms: the
normwise bound of their arithmetic), and for byte-identical reruns; the plan's shape is checked
where a rule fixes it."""
import math
import random
import sys

import pytest

import oracle
from helpers import check_against_oracle, inputs, ptr, run_gpu, to_dev

pytestmark = pytest.mark.gpu

ALPHA, BETA = complex(0.7, -0.3), complex(0.2, 0.9)
OPS = ["NN", "NT", "TN", "TT", "NC", "CN", "CC", "TC", "CT"]
# normwise error bound against FP64 (uniform[-1,1] inputs, k up to ~1e4)
BOUND = {"bf16x3": 6e-5, "bf16x6": 4e-6, "fp16": 4e-3}


def _split(rng, mnl):
    """m, n with floor(log2(m n)) == mnl and a random aspect ratio."""
    mn = rng.randint(2 ** mnl, 2 ** (mnl + 1) - 1)
    m = max(1, int(round(math.sqrt(mn * rng.uniform(0.3, 3.0)))))
    n = max(1, mn // m)
    while int(math.floor(math.log2(m * n))) != mnl:
        n += 1 if m * n < 2 ** mnl else -1
    return m, n


def _case(h, ops, m, n, k, seed, cfg=None):
    ta, tb = ops[0], ops[1]
    print("case %s m=%d n=%d k=%d cfg=%s" % (ops, m, n, k, cfg))
    sys.stdout.flush()
    lda = (m if ta == "N" else k) + 1
    ldb = (k if tb == "N" else n) + 1
    ldc = m + 2
    A, B, C = inputs(ta, tb, m, n, k, lda, ldb, ldc, seed)
    dA, dB = to_dev(A), to_dev(B)
    st, out = run_gpu(h, ta, tb, m, n, k, ALPHA, A, lda, B, ldb, BETA, C, ldc, cfg, dA, dB)
    assert st == 0, st
    plan = h.last_plan(sync=True)
    st2, out2 = run_gpu(h, ta, tb, m, n, k, ALPHA, A, lda, B, ldb, BETA, C, ldc, cfg, dA, dB)
    same = st2 == 0 and out2.tobytes() == out.tobytes()
    assert same, "rerun differs"
    return plan, out, (A, B, C, lda, ldb, ldc, dA, dB)


def test_fp32_small_outputs_long_k(gpu):
    """FP32, outputs of 40..400 x 40..400 with k of 2048..8192: the packed path with the
    K split the planner picks for classes the table does not list (or the table's plan)."""
    rng = random.Random(31)
    for i in range(12):
        m, n, k = rng.randint(40, 400), rng.randint(40, 400), rng.randint(2048, 8192)
        ops = rng.choice(OPS)
        plan, out, (A, B, C, lda, ldb, ldc, _, _) = _case(gpu, ops, m, n, k, 7100 + i)
        gold = oracle.cgemm(ops[0], ops[1], m, n, k, ALPHA, A, lda, B, ldb, BETA, C, ldc)
        check_against_oracle(out, gold, C, m, n, ldc, "fp32 %s %dx%dx%d plan %s" % (ops, m, n, k, plan))


def _neighbour_classes():
    """Classes (floor(log2(m n)), floor(log2(k))) the table does not list whose two k neighbours
    it lists as the one-launch kernel without k slices (tile 16 or 32, parts 1)."""
    from cgemm3m import autotune

    table = {key: val for key, _, val in autotune.read_single_entries()}

    def one(e):
        return e is not None and e[0] in (16, 32) and e[1] == 1

    return [(a, b) for a in range(0, 40) for b in range(1, 40)
            if (a, b) not in table and one(table.get((a, b - 1))) and one(table.get((a, b + 1)))]


def test_fp32_one_launch_between_one_launch_classes(gpu):
    """FP32 shapes in classes the table does not list whose two k neighbours take the one-launch
    kernel without k slices: the plan is the one-launch kernel, unsliced, and meets the oracle."""
    classes = _neighbour_classes()
    assert classes
    rng = random.Random(32)
    for i, (mnl, kl) in enumerate(classes):
        m, n = _split(rng, mnl)
        k = rng.randint(2 ** kl, 2 ** (kl + 1) - 1)
        ops = rng.choice(OPS)
        plan, out, (A, B, C, lda, ldb, ldc, _, _) = _case(gpu, ops, m, n, k, 7200 + i)
        assert plan["path"] == 5 and plan["parts"] == 1, plan
        gold = oracle.cgemm(ops[0], ops[1], m, n, k, ALPHA, A, lda, B, ldb, BETA, C, ldc)
        check_against_oracle(out, gold, C, m, n, ldc, "fp32 one-launch %s %dx%dx%d" % (ops, m, n, k))


@pytest.mark.parametrize("prec", ["bf16x3", "bf16x6", "fp16"])
def test_16bit_small_outputs_split_k(gpu, prec):
    """16-bit forms on small outputs with long k, the planner's K split: within the normwise
    bound against FP64, reruns byte-identical; the BF16 separate products' split keeps its
    workgroups within one per CU."""
    import cgemm3m

    rng = random.Random(33 + len(prec))
    for i in range(6):
        m, n, k = rng.randint(100, 1300), rng.randint(100, 1300), rng.randint(1000, 9000)
        ops = rng.choice(OPS)
        cfg = cgemm3m.config(prec)
        plan, out, (A, B, C, lda, ldb, ldc, dA, dB) = _case(gpu, ops, m, n, k, 7300 + i, cfg)
        if prec != "fp16" and plan["path"] == 4 and plan["parts"] > 1:
            tile = plan["tile"]
            wgs = -(-m // tile) * -(-n // tile) * (3 if tile == 256 else 1) * plan["parts"]
            assert wgs <= 256, (plan, wgs)
        dC0, dOut = to_dev(C), to_dev(out)
        res = cgemm3m.residual(gpu, ops[0], ops[1], m, n, k, ALPHA, ptr(dA), lda, ptr(dB), ldb, BETA, ptr(dC0),
                               ptr(dOut), ldc)
        print("plan", plan, "err", res)
        assert res["normwise"] <= BOUND[prec], (prec, ops, m, n, k, plan, res)
