"""The generated functional case list (no GPU): its size, categories and coverage, and that it
is the same list on every run."""
import collections

from suite_cases import OPS, chunks, suite_cases


def test_suite_is_deterministic_and_sized():
    a, b = suite_cases(), suite_cases()
    assert a == b
    cats = collections.Counter(c["cat"] for c in a)
    assert cats == {"square": 108, "rect": 81, "sweep": 663, "scalars": 36, "ld": 16, "complex_scalars": 32}
    assert sum(len(ch) for ch in chunks(a)) == len(a)
    assert len({c["seed"] for c in a}) == len(a)


def test_suite_covers_ops_scalars_and_layouts():
    cases = suite_cases()
    assert {c["ta"] + c["tb"] for c in cases} == set(OPS)
    for cat in ("square", "rect", "sweep", "scalars"):
        assert {c["ta"] + c["tb"] for c in cases if c["cat"] == cat} == set(OPS), cat
    assert any(c["alpha"] == 0 for c in cases)
    assert any(c["beta"] == 0 for c in cases) and any(c["beta"] == 1j for c in cases)
    for c in cases:
        rows_a = c["m"] if c["ta"] == "N" else c["k"]
        rows_b = c["k"] if c["tb"] == "N" else c["n"]
        assert c["lda"] >= max(1, rows_a) and c["ldb"] >= max(1, rows_b) and c["ldc"] >= max(1, c["m"])
        assert min(c["m"], c["n"], c["k"]) >= 1
    assert any(c["ldc"] > c["m"] for c in cases)  # padding rows to keep untouched
    assert max(max(c["m"], c["n"], c["k"]) for c in cases) == 2048
