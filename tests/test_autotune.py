"""The autotuner's host logic (no GPU): plan choice with its noise margin, per-class picks,
its own generated shape set, the table it writes, and the kernel-trace reader."""
import csv
import math
import re

from cgemm3m import autotune as at


def test_bin_of_is_floor_log2():
    assert at.bin_of(256, 256, 256) == (16, 8)
    assert at.bin_of(581, 782, 175) == (18, 7)
    assert at.bin_of(1, 1, 1) == (0, 0)
    assert at.bin_of(8192, 8192, 8191) == (26, 12)


def test_choose_prefers_fewer_partitions_then_larger_tile_within_margin():
    # within 2% of the fastest: fewest partitions, then the larger tile
    assert at.choose({(64, 2): 1.00, (128, 1): 1.015, (64, 1): 1.019}) == (128, 1)
    # outside the margin the fastest wins
    assert at.choose({(64, 2): 1.00, (128, 1): 1.03}) == (64, 2)
    assert at.choose({(16, 1): 2.0}) == (16, 1)


def test_pick_bins_minimises_geometric_mean_of_time_over_best():
    # two shapes in one class: (32, 1) is best for one, (16, 1) for the other; the geometric
    # mean of time / best picks the plan that loses least overall
    measured = [
        (("N", "N", 512, 512, 200), {(32, 1): 1.0, (16, 1): 1.5}),
        (("T", "N", 500, 600, 150), {(32, 1): 1.1, (16, 1): 1.0}),
    ]
    assert at.pick_bins(measured) == [((18, 7), 0, (32, 1))]


def test_generated_shapes_are_seeded_and_cover_the_range():
    a, b = at.generated_shapes(), at.generated_shapes()
    assert a == b and len(a) == 320
    assert {s[0] + s[1] for s in a} == {x + y for x in "NTC" for y in "NTC"}
    dims = [d for s in a for d in s[2:5]]
    assert min(dims) == 1 or min(dims) <= 2
    assert max(dims) <= 4096 and max(dims) >= 3500
    for ta, tb, m, n, k, lda, ldb, ldc in a:
        assert lda >= (m if ta == "N" else k) and ldb >= (k if tb == "N" else n) and ldc >= m
    # at least a few shapes in each decade of m*n*k, from tiny to 4096^3-sized
    mnk = [math.log10(s[2] * s[3] * s[4]) for s in a]
    assert sum(x < 3 for x in mnk) >= 3 and sum(x > 10 for x in mnk) >= 3


def test_write_table_round_trip(tmp_path):
    entries = [((18, 7), 0, (32, 1)), ((12, 6), 1, (16, 1)), ((20, 9), 1, (128, 2))]
    path = tmp_path / "table.h"
    at.write_table(entries, str(path))
    text = path.read_text()
    assert "{18, 7, 0, 32, 1}," in text and "{12, 6, 1, 16, 1}," in text
    assert "kTunedShapes" not in text
    assert sorted(at.read_grouped_entries(str(path))) == [((12, 6), 1, (16, 1)), ((20, 9), 1, (128, 2))]


def test_installed_table_is_consistent():
    text = open(at.TABLE).read()
    assert "kTunedShapes" not in text  # classes only: no plan for any single shape
    classes = text[text.index("kTunedTable[]"):]
    bins = {}
    for a, b, g, t, p in re.findall(r"\{(\d+), (\d+), (\d+), (\d+), (\d+)\}", classes):
        if int(t):
            key = (int(a), int(b), int(g))
            assert key not in bins, key
            assert int(t) in (16, 32, 64, 128) and 1 <= int(p) <= 16
            bins[key] = (int(t), int(p))


def test_read_trace_groups_candidates_and_skips_the_pool_clear(tmp_path):
    calls = at.CALLS
    rows, t = [], 0

    def kernel(name, dur):
        nonlocal t
        rows.append({"Kernel_Name": name, "Start_Timestamp": str(t), "End_Timestamp": str(t + dur)})
        t += dur + 10

    # candidate (32, 2): one-time pool clear in its first call, then one kernel per call
    kernel("at::native::fill", 5)
    kernel("c3m::fill_u64_kernel(unsigned long long*, long long)", 3000)
    for _ in range(calls):
        kernel("void c3m::gemm3m_small_kernel<4>(c3m::SmallArgs, int, int)", 20000)
    # candidate (64, 1): pack + product per call
    kernel("at::native::fill", 5)
    for i in range(calls):
        kernel("c3m::pack_kernel(c3m::PackArgs, c3m::PackArgs, int)", 7000)
        kernel("void c3m::gemm3m_f32_kernel<64, 64, false, 2, 2>(c3m::GemmArgs)", 17000 + 1000 * (i < 2))
    kernel("at::native::fill", 5)
    trace = tmp_path / "trace.csv"
    with open(trace, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        w.writerows(rows)
    manifest = tmp_path / "manifest.json"
    manifest.write_text('{"calls": %d, "entries": [{"shape": ["T", "N", 581, 782, 175], "cands": [[32, 2], [64, 1]]}]}'
                        % calls)
    (shape, res), = at.read_trace(str(trace), str(manifest))
    assert shape == ("T", "N", 581, 782, 175)
    # ns -> ms, warm-up calls (the first two) left out
    assert abs(res[(32, 2)] - 0.020) < 1e-12
    assert abs(res[(64, 1)] - 0.024) < 1e-12
