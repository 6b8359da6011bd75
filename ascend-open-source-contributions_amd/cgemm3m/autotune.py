"""Autotuner for the Cgemm3m planner: measures plan candidates per shape class on the GPU
and writes csrc/tuned_table.h (compiled into the library).

A class is the bin (floor(log2(m*n)), floor(log2(k))). One-problem candidates: output tile 64
or 128 with 1-16 K partitions, and the one-launch small-problem kernel with four waves per
16x16 output tile ("tile 32") or one ("tile 16"), with 1-16 k slices per output tile. Grouped
classes are timed as one aclblasCgemm3mGroupedBatchedEx call of identical problems with the
member plan pinned. The table holds classes only (no entry for any single shape) and is
deterministic input to the planner: the same arguments always get the same plan.

One-problem timing, as the reference measures its performance cases (ops-blas/test/gemm/
cgemm3m/README.md:95-115): the sum of a call's kernels' device durations from a rocprofv3
kernel trace, launch gaps excluded. Two steps on the GPU:

  rocprofv3 --kernel-trace --output-format csv -d OUT -o run -- python3 -m cgemm3m.autotune run MANIFEST
  python3 -m cgemm3m.autotune pick OUT/run_kernel_trace.csv MANIFEST [--write]

`run` separates candidates with one non-library kernel (a torch fill) and writes MANIFEST
(shapes and candidate order); `pick` chooses, per class, the candidate with the smallest
geometric mean of time / (that shape's best time) over the class's shapes, and also times the
grouped classes with events. Without a profiler, `python -m cgemm3m.autotune events` times
every candidate with events around the call (launch gaps included); `grouped` re-times only
the grouped classes and keeps the installed table's one-problem classes.

The shapes: a few squares and tall/flat problems, plus this project's own seeded set of
generated shapes (`generated_shapes`: every dimension 1..4096, all nine op pairs, some padded
leading dimensions). The reference's performance suite is not among them: `pf_suite.py` at the
repository root measures it as a held-out set.
"""
import argparse
import json
import math
import os
import random
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.normpath(os.path.join(HERE, "..", "csrc", "tuned_table.h"))
CALLS = 6  # per candidate; the first two are warm-up

SINGLE_SHAPES = [
    (256, 256, 256), (512, 512, 512), (1024, 1024, 1024), (2048, 2048, 2048), (4096, 4096, 4096),
    (8192, 8192, 8192), (1000, 1500, 2048), (256, 256, 65536), (256, 256, 8192), (512, 512, 16384),
    (128, 128, 4096), (2048, 2048, 256), (4096, 4096, 512), (64, 64, 1024),
    (64, 64, 64), (128, 128, 128), (128, 128, 512), (384, 384, 384), (512, 512, 128), (768, 768, 768),
    (1024, 1024, 256), (1024, 1024, 512),
    # rectangular mid sizes (every order of 1024, 2048, 4096: the products of a 2048 x 1024 by
    # 1024 x 4096 forward and its two gradient products)
    (1024, 2048, 4096), (1024, 4096, 2048), (2048, 1024, 4096), (2048, 4096, 1024), (4096, 1024, 2048),
    (4096, 2048, 1024),
]
# square members (even log2 m*n classes) and m x 2m members (the odd classes between them)
GROUPED_SHAPES = [(m, m * r, k) for m in (32, 64, 128, 256, 512) for r in (1, 2) if m * r <= 512
                  for k in (64, 256, 1024, 2048)]


def generated_shapes(count=320, seed=20261204):
    """This project's tuning set: `count` shapes with every dimension in 1..4096 (half drawn
    log-uniformly, half uniformly, so both the small sizes and the mid sizes are dense), op
    pairs uniform over the nine, and the stored rows as leading dimensions plus a pad of 1-7
    for about a quarter of them. Seeded: the same list on every machine."""
    rng = random.Random(seed)

    def dim():
        if rng.random() < 0.5:
            return max(1, min(4096, int(round(2.0 ** rng.uniform(0.0, 12.0)))))
        return rng.randint(1, 4096)

    def pad():
        return rng.randint(1, 7) if rng.random() < 0.25 else 0

    out = []
    for _ in range(count):
        ta, tb = rng.choice("NTC"), rng.choice("NTC")
        m, n, k = dim(), dim(), dim()
        lda = (m if ta == "N" else k) + pad()
        ldb = (k if tb == "N" else n) + pad()
        out.append((ta, tb, m, n, k, lda, ldb, m + pad()))
    return out


def bin_of(m, n, k):
    return int(math.floor(math.log2(m * n))), int(math.floor(math.log2(k)))


def candidates(m, n, k):
    """(tile, parts) plans worth timing for one problem (tile 32 / 16: the one-launch kernel with
    four waves / one wave per output tile, parts = k slices)."""
    steps = (k + 31) // 32 * 2          # 16-deep stages of the packed kernels
    chunks = (k + 15) // 16             # 16-deep k chunks of the small kernel
    out = [(tile, parts) for tile in (64, 128) for parts in (1, 2, 4, 8, 16)
           if parts == 1 or steps // parts >= 2]
    if m * n <= 2048 * 2048:
        out += [(32, s) for s in (1, 2, 4, 8, 16) if s == 1 or chunks // s >= 4]
        out += [(16, s) for s in (1, 2, 4) if s == 1 or chunks // s >= 2]
    return out


def _config(cgemm3m, tile, parts):
    if tile in (16, 32):
        return cgemm3m.config("fp32", variant=5, k_split=parts, tile=16 if tile == 16 else 0)
    return cgemm3m.config("fp32", tile=tile, k_split=parts)


def shape_list(which):
    """[(ta, tb, m, n, k, lda, ldb, ldc)] for 'builtin', 'gen' or 'all'."""
    out = []
    if which in ("builtin", "all"):
        out += [("N", "N", m, n, k, m, k, m) for (m, n, k) in SINGLE_SHAPES]
    if which in ("gen", "all"):
        out += generated_shapes()
    return out


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


def _operands(torch, ta, tb, m, n, k, lda, ldb, ldc):
    A = (torch.rand(2 * lda * (k if ta == "N" else m), device="cuda") * 2 - 1).view(torch.complex64)
    B = (torch.rand(2 * ldb * (n if tb == "N" else k), device="cuda") * 2 - 1).view(torch.complex64)
    C = torch.zeros(ldc * n, dtype=torch.complex64, device="cuda")
    return A, B, C


def tune_single_events(h, torch, shapes, reps):
    """Every candidate timed with events around the call (median of `reps`)."""
    import cgemm3m

    timed = _timer(torch)
    out = []
    P = cgemm3m.torch_ptr
    for (ta, tb, m, n, k, lda, ldb, ldc) in shapes:
        A, B, C = _operands(torch, ta, tb, m, n, k, lda, ldb, ldc)
        res = {}
        for tile, parts in candidates(m, n, k):
            cfg = _config(cgemm3m, tile, parts)
            res[(tile, parts)] = timed(lambda: cgemm3m.cgemm3m(h, ta, tb, m, n, k, 1.0, P(A), lda, P(B), ldb, 0.0,
                                                               P(C), ldc, cfg), reps)
        out.append(((ta, tb, m, n, k), res))
        print("single %s%s %5d x %5d x %6d  best %s" % (ta, tb, m, n, k, min(res, key=res.get)), flush=True)
        del A, B, C
    return out


def run_trace(h, torch, shapes, manifest):
    """Every candidate CALLS times, candidates separated by a torch fill (for `pick`)."""
    import cgemm3m

    P = cgemm3m.torch_ptr
    marker = torch.zeros(64, device="cuda")
    entries = []
    for idx, (ta, tb, m, n, k, lda, ldb, ldc) in enumerate(shapes):
        A, B, C = _operands(torch, ta, tb, m, n, k, lda, ldb, ldc)
        cands = candidates(m, n, k)
        for tile, parts in cands:
            cfg = _config(cgemm3m, tile, parts)
            torch.cuda.synchronize()
            marker.fill_(float(idx))
            for _ in range(CALLS):
                cgemm3m.cgemm3m(h, ta, tb, m, n, k, 1.0, P(A), lda, P(B), ldb, 0.0, P(C), ldc, cfg)
                torch.cuda.synchronize()
        entries.append({"shape": [ta, tb, m, n, k], "cands": [list(c) for c in cands]})
        if idx % 20 == 19:
            print("autotune run: %d / %d shapes" % (idx + 1, len(shapes)), flush=True)
        del A, B, C
    torch.cuda.synchronize()
    marker.fill_(-1.0)
    torch.cuda.synchronize()
    with open(manifest, "w") as f:
        json.dump({"calls": CALLS, "entries": entries}, f)


def read_trace(trace, manifest):
    """[((ta, tb, m, n, k), {(tile, parts): mean kernel-time sum per call in ms})] from `run`."""
    import csv

    with open(manifest) as f:
        man = json.load(f)
    calls = man["calls"]
    groups, cur = [], None
    with open(trace) as f:
        rows = sorted(csv.DictReader(f), key=lambda r: int(r["Start_Timestamp"]))
    for r in rows:
        if "c3m::" in r["Kernel_Name"]:
            if cur is None:
                cur = []
                groups.append(cur)
            cur.append((r["Kernel_Name"], (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6))
        else:
            cur = None
    want = sum(len(e["cands"]) for e in man["entries"])
    if len(groups) != want:
        raise SystemExit("trace holds %d candidate groups, the manifest %d" % (len(groups), want))
    out, g = [], 0
    for e in man["entries"]:
        res = {}
        for c in e["cands"]:
            ks = groups[g]
            g += 1
            if len(ks) % calls == 1 and "fill_u64_kernel" in ks[0][0]:
                ks = ks[1:]  # the handle's one-time clear of its arrival-counter pool
            ks = [t for _, t in ks]
            if len(ks) % calls:
                raise SystemExit("shape %s candidate %s: %d kernels for %d calls" % (e["shape"], c, len(ks), calls))
            per = len(ks) // calls
            times = [sum(ks[i * per:(i + 1) * per]) for i in range(2, calls)]
            res[tuple(c)] = sum(times) / len(times)
        out.append((tuple(e["shape"]), res))
    return out


def pick_bins(measured):
    """Per bin, the candidate timed on every shape of the bin with the smallest geometric mean
    of time / best time for that shape (ties: fewer partitions, then the larger tile)."""
    bins = {}
    for shape, res in measured:
        bins.setdefault(bin_of(*shape[2:5]), []).append(res)
    out = []
    for b, members in sorted(bins.items()):
        common = set(members[0])
        for res in members[1:]:
            common &= set(res)
        score = {}
        for c in common:
            score[c] = math.exp(sum(math.log(res[c] / min(res.values())) for res in members) / len(members))
        out.append((b, 0, choose(score)))
    return out


def read_single_entries(path=TABLE):
    """The one-problem class entries of an existing table: [((mnLog2, kLog2), 0, (tile, parts))]."""
    text = open(path).read()
    body = text[text.index("kTunedTable[]"):]
    body = body[:body.index("};")]
    out = []
    for a, b, g, t, p in re.findall(r"\{(\d+), (\d+), (\d+), (\d+), (\d+)\}", body):
        if int(g) == 0 and int(t) != 0:
            out.append(((int(a), int(b)), 0, (int(t), int(p))))
    return out


def read_grouped_entries(path=TABLE):
    """The grouped class entries of an existing table: [((mnLog2, kLog2), 1, (tile, parts))]."""

    text = open(path).read()
    body = text[text.index("kTunedTable[]"):]
    body = body[:body.index("};")]
    out = []
    for a, b, g, t, p in re.findall(r"\{(\d+), (\d+), (\d+), (\d+), (\d+)\}", body):
        if int(g) == 1 and int(t) != 0:
            out.append(((int(a), int(b)), 1, (int(t), int(p))))
    return out


OP_PAIRS = [(a, b) for a in "NTC" for b in "NTC"]


def tune_grouped(h, torch, shapes, reps, count=63):
    """Grouped classes: one grouped call of `count` problems of one shape per candidate, their
    op pairs cycling through all nine (the persistent kernel reads operands as stored, so its
    speed depends on their layout), with the member plan pinned through
    aclblasCgemm3mGroupedBatchedEx."""
    import cgemm3m

    timed = _timer(torch)
    out = []
    for (m, n, k) in shapes:
        keep = []
        probs = []
        for i in range(count):
            ta, tb = OP_PAIRS[i % len(OP_PAIRS)]
            a = torch.rand(2 * m * k, device="cuda").view(torch.complex64)
            b = torch.rand(2 * k * n, device="cuda").view(torch.complex64)
            c = torch.zeros(m * n, dtype=torch.complex64, device="cuda")
            keep += [a, b, c]
            probs.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=1.0, A=cgemm3m.torch_ptr(a),
                              lda=m if ta == "N" else k, B=cgemm3m.torch_ptr(b), ldb=k if tb == "N" else n,
                              beta=0.0, C=cgemm3m.torch_ptr(c), ldc=m))
        res = {}
        for tile in (64, 128):
            for parts in (1, 2, 4):
                if parts > 1 and k // parts < 256:
                    continue
                cfg = cgemm3m.config("fp32", tile=tile, k_split=parts)
                res[(tile, parts)] = timed(lambda: cgemm3m.grouped(h, probs, cfg=cfg), reps)
        if m * n <= 512 * 512:  # the one-launch kernel, four waves / one wave per tile, unsliced
            for tile in (32, 16):
                cfg = cgemm3m.config("fp32", variant=5, tile=16 if tile == 16 else 0, k_split=1)
                res[(tile, 1)] = timed(lambda: cgemm3m.grouped(h, probs, cfg=cfg), reps)
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
        "// (tile 32 / 16: the one-launch small-problem kernel, four waves / one wave per 16x16 tile;",
        "// parts = k slices). Classes only, tuned on the tuner's own generated shapes; bins not",
        "// listed fall back to the planner's occupancy rule.",
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
    ap.add_argument("mode", choices=["run", "pick", "events", "grouped"])
    ap.add_argument("paths", nargs="*", help="run: MANIFEST; pick: TRACE MANIFEST; grouped: none (re-times "
                                             "the grouped classes, keeps the installed one-problem ones)")
    ap.add_argument("--shapes", default="all", choices=["builtin", "gen", "all"])
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--write", action="store_true")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--keep-grouped", action="store_true",
                    help="pick: reuse the grouped entries of the installed table instead of timing them "
                         "(no GPU needed)")
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    if args.mode == "pick" and len(args.paths) != 2:
        raise SystemExit("pick needs TRACE MANIFEST")
    if args.mode == "run" and len(args.paths) != 1:
        raise SystemExit("run needs MANIFEST")
    import torch

    import cgemm3m

    h = cgemm3m.Handle()
    if args.mode == "grouped":
        grouped = GROUPED_SHAPES[::3] if args.quick else GROUPED_SHAPES
        entries = read_single_entries() + tune_grouped(h, torch, grouped, args.reps)
        if args.write:
            print("wrote", write_table(entries, args.out))
        h.close()
        return
    shapes = shape_list(args.shapes)
    if args.quick:
        shapes = shapes[::10]
    if args.mode == "run":
        run_trace(h, torch, shapes, args.paths[0])
        h.close()
        return
    measured = read_trace(args.paths[0], args.paths[1]) if args.mode == "pick" else \
        tune_single_events(h, torch, shapes, args.reps)
    single = pick_bins(measured)
    for (b, _, best) in single:
        print("bin mn 2^%d k 2^%d: tile %d parts %d" % (b[0], b[1], best[0], best[1]))
    if args.keep_grouped:
        grouped_entries = read_grouped_entries()
    else:
        grouped = GROUPED_SHAPES[::3] if args.quick else GROUPED_SHAPES
        grouped_entries = tune_grouped(h, torch, grouped, args.reps)
    if args.write:
        print("wrote", write_table(single + grouped_entries, args.out))
    h.close()


if __name__ == "__main__":
    main()
