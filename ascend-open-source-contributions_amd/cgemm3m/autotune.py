"""Autotuner for the Cgemm3m planner: measures output tile (64/128) x K partitions per shape
class on the GPU and writes csrc/tuned_table.h (compiled into the library).

One-problem classes are timed with aclblasCgemm3mEx and pinned plans; grouped classes are
timed as one aclblasCgemm3mGroupedBatchedEx call of identical problems with the plan pinned. A class is the
bin (floor(log2(m*n)), floor(log2(k))). Each candidate's time is the median of `reps`
event-timed calls after one warm-up. The table is deterministic input to the planner: the
same arguments always get the same plan.

Usage (GPU):  python -m cgemm3m.autotune [--quick] [--write]
"""
import argparse
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.normpath(os.path.join(HERE, "..", "csrc", "tuned_table.h"))

SINGLE_SHAPES = [
    (256, 256, 256), (512, 512, 512), (1024, 1024, 1024), (2048, 2048, 2048), (4096, 4096, 4096),
    (8192, 8192, 8192), (1000, 1500, 2048), (256, 256, 65536), (256, 256, 8192), (512, 512, 16384),
    (128, 128, 4096), (2048, 2048, 256), (4096, 4096, 512), (64, 64, 1024),
    (64, 64, 64), (128, 128, 128), (128, 128, 512), (384, 384, 384), (512, 512, 128), (768, 768, 768),
    (1024, 1024, 256), (1024, 1024, 512),
]
GROUPED_SHAPES = [(m, m, k) for m in (32, 64, 128, 256, 512) for k in (64, 256, 1024, 2048)]


def bin_of(m, n, k):
    return int(math.floor(math.log2(m * n))), int(math.floor(math.log2(k)))


def choose(res, margin=0.02):
    """Best (tile, parts) of {(tile, parts): ms}: among the candidates within `margin` of the
    fastest, the one with the fewest K partitions (less workspace, no merge pass), then the
    larger tile, so timing noise does not turn a tie into a K split."""
    best = min(res.values())
    close = [c for c, t in res.items() if t <= best * (1.0 + margin)]
    return min(close, key=lambda c: (c[1], -c[0]))


def _timer(torch):
    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b))
        ts.sort()
        return ts[len(ts) // 2]
    return timed


def tune_single(h, torch, shapes, reps):
    import cgemm3m

    timed = _timer(torch)
    out = []
    for (m, n, k) in shapes:
        A = torch.rand(2 * m * k, device="cuda").view(torch.complex64)
        B = torch.rand(2 * k * n, device="cuda").view(torch.complex64)
        C = torch.zeros(m * n, dtype=torch.complex64, device="cuda")
        P = cgemm3m.torch_ptr
        res = {}
        cands = [(tile, parts) for tile in (64, 128) for parts in (1, 2, 4, 8, 16) if parts == 1 or k // parts >= 256]
        if m * n <= 2048 * 2048:
            cands.append((32, 1))  # the one-launch small-problem kernel
        for tile, parts in cands:
            if tile == 32:
                cfg = cgemm3m.config("fp32", variant=5)
            else:
                cfg = cgemm3m.config("fp32", tile=tile, k_split=parts)
            res[(tile, parts)] = timed(lambda: cgemm3m.cgemm3m(h, "N", "N", m, n, k, 1.0, P(A), m, P(B), k,
                                                               0.0, P(C), m, cfg), reps)
        best = choose(res)
        print("single %5d x %5d x %6d  best tile %d parts %2d  %.3f ms  (all: %s)" % (
            m, n, k, best[0], best[1], res[best], {"%d/%d" % kk: round(v, 3) for kk, v in res.items()}),
            flush=True)
        out.append((bin_of(m, n, k), 0, best))
        del A, B, C
    return out


def tune_grouped(h, torch, shapes, reps, count=64):
    """Grouped classes: one grouped call of `count` identical problems per candidate, with
    the member plan pinned through aclblasCgemm3mGroupedBatchedEx."""
    import cgemm3m

    timed = _timer(torch)
    out = []
    for (m, n, k) in shapes:
        keep = []
        probs = []
        for _ in range(count):
            a = torch.rand(2 * m * k, device="cuda").view(torch.complex64)
            b = torch.rand(2 * k * n, device="cuda").view(torch.complex64)
            c = torch.zeros(m * n, dtype=torch.complex64, device="cuda")
            keep += [a, b, c]
            probs.append(dict(ta="N", tb="N", m=m, n=n, k=k, alpha=1.0, A=cgemm3m.torch_ptr(a), lda=m,
                              B=cgemm3m.torch_ptr(b), ldb=k, beta=0.0, C=cgemm3m.torch_ptr(c), ldc=m))
        res = {}
        for tile in (64, 128):
            for parts in (1, 2, 4):
                if parts > 1 and k // parts < 256:
                    continue
                cfg = cgemm3m.config("fp32", tile=tile, k_split=parts)
                res[(tile, parts)] = timed(lambda: cgemm3m.grouped(h, probs, cfg=cfg), reps)
        best = choose(res)
        print("grouped %4d x %4d x %5d  best tile %d parts %d  (all: %s)" % (
            m, n, k, best[0], best[1], {"%d/%d" % kk: round(v, 3) for kk, v in res.items()}), flush=True)
        out.append((bin_of(m, n, k), 1, best))
        del keep
    return out


def write_table(entries, path=TABLE):
    seen = {}
    for (mnl, kl), grouped, (tile, parts) in entries:
        seen[(mnl, kl, grouped)] = (tile, parts)
    lines = [
        "// tuned_table.h -- plan table written by cgemm3m/autotune.py from MI355X measurements.",
        "// Key: (floor(log2(m*n)), floor(log2(k)), grouped); value: output tile edge and K partitions",
        "// (tile 32: the one-launch small-problem kernel).",
        "// Bins not listed fall back to the planner's occupancy rule.",
        "#pragma once",
        "",
        "namespace c3m {",
        "",
        "struct TunedEntry {",
        "    short mnLog2, kLog2, grouped, tile, parts;",
        "};",
        "",
        "static const TunedEntry kTunedTable[] = {",
    ]
    for (mnl, kl, g), (tile, parts) in sorted(seen.items()):
        lines.append("    {%d, %d, %d, %d, %d}," % (mnl, kl, g, tile, parts))
    lines += ["    {0, 0, 0, 0, 0},  // sentinel", "};", "", "}  // namespace c3m", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))
    return path


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--write", action="store_true")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=TABLE)
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    import torch

    import cgemm3m

    h = cgemm3m.Handle()
    single = SINGLE_SHAPES[:6] if args.quick else SINGLE_SHAPES
    grouped = GROUPED_SHAPES[::3] if args.quick else GROUPED_SHAPES
    entries = tune_single(h, torch, single, args.reps) + tune_grouped(h, torch, grouped, args.reps)
    if args.write:
        print("wrote", write_table(entries, args.out))
    h.close()


if __name__ == "__main__":
    main()
