#!/usr/bin/env python3
"""Cgemm3m throughput on one MI355X: complex TFLOP/s (8mnk) at m = n = k = 8192.

One step = one aclblasCgemm3m call (pack -> three real products -> alpha/beta epilogue ->
rounding correction) over 8192^3 complex64 operands already resident in HBM, with
alpha = (1, 0.5), beta = (0.5, -0.25), op = (N, N). `value` is the plain FP32-MFMA path;
the split-precision forms (bf16x3, bf16x6, fp16 block-scaled) and the (C, T) ops are
reported beside it, each with its max/mean relative error against an FP64 GEMM of the same
inputs computed on the GPU, and whether two runs are byte-identical.

Prints progress on stderr and ONE JSON line on stdout.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "ascend-open-source-contributions_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "complex TFLOP/s (8mnk) Cgemm3m 8192^3: bf16x3 vs FP32-MFMA, rel err vs FP64"
PEAK = {"fp32": 157.3, "bf16x3": 2516.6, "bf16x6": 2516.6, "fp16": 2516.6}   # dense TFLOP/s
MFMA_FLOPS_PER_MNK = {"fp32": 6, "bf16x3": 18, "bf16x6": 36, "fp16": 6}       # algorithmic


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--variants", default="fp32,bf16x3,bf16x6,fp16")
    ap.add_argument("--ops", default="NN,CT")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-errors", action="store_true")
    ap.add_argument("--no-grouped", action="store_true")
    ap.add_argument("--no-vendor", action="store_true", help="skip the torch.matmul complex64 comparison")
    ap.add_argument("--no-autograd", action="store_true", help="skip the autograd (forward + backward) timing")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--profile-run", action="store_true",
                    help="for rocprofv3 counter passes: 3 NN calls of the first --variants form (fp32 "
                         "by default), then one device copy of exactly 1 GiB (calibrates FETCH_SIZE/"
                         "WRITE_SIZE), nothing timed")
    ap.add_argument("--profile-variant", type=int, default=0,
                    help="with --profile-run: the kernel variant to pin (aclblasCgemm3mConfig.variant; 0 = the "
                         "planner's choice)")
    ap.add_argument("--grouped-only", action="store_true",
                    help="for kernel traces: only the 512-problem grouped batch (one call, its hipGraph "
                         "replay, the per-problem loop); prints its figures as one JSON line. With "
                         "--profile-run: three grouped calls, then the 1 GiB copy, nothing timed")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path here (across GPUs the "
              "operator would shard C by output blocks, each GPU holding its rows of op(A) and columns "
              "of op(B), with no exchange)", file=sys.stderr)
        sys.exit(2)

    import torch

    import cgemm3m

    m = n = k = args.size
    dev = torch.device("cuda:0")
    if args.grouped_only:
        hg = cgemm3m.Handle(stream=torch.cuda.current_stream())
        if args.profile_run:
            probs, keep, _ = grouped_problems(torch, cgemm3m, dev)
            for _ in range(3):
                cgemm3m.grouped(hg, probs)
            src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
            dst = torch.empty_like(src)
            dst.copy_(src)
            torch.cuda.synchronize()
            log("profile run done: 3 grouped calls of 512 problems, then a 1 GiB copy")
        else:
            print(json.dumps({"grouped_512": grouped_bench(hg, torch, cgemm3m, dev)}), flush=True)
        hg.close()
        return
    gen = torch.Generator(device=dev)
    gen.manual_seed(20261203)

    def uniform_c64(count):
        return (torch.rand(2 * count, generator=gen, device=dev) * 2 - 1).view(torch.complex64)

    A = uniform_c64(m * k)
    B = uniform_c64(k * n)
    C0 = uniform_c64(m * n)
    alpha, beta = complex(1.0, 0.5), complex(0.5, -0.25)
    stream = torch.cuda.current_stream()
    h = cgemm3m.Handle(stream=stream)
    P = cgemm3m.torch_ptr

    def call(cfg, ops, Cbuf, Aop=None):
        ta, tb = ops[0], ops[1]
        Ause = A if Aop is None else Aop
        # square operands: every leading dimension is m
        return cgemm3m.cgemm3m(h, ta, tb, m, n, k, alpha, P(Ause), m, P(B), m, beta, P(Cbuf), m, cfg, check=False)

    if args.profile_run:
        # the first of --variants (fp32 unless --variants is given)
        form = args.variants.split(",")[0] if args.variants != ap.get_default("variants") else "fp32"
        C = C0.clone()
        for _ in range(3):
            st = call(cgemm3m.config(form, variant=args.profile_variant), "NN", C)
            if st != 0:
                log("profile run: %s variant %d: status %s" % (form, args.profile_variant, cgemm3m._lib.STATUS.get(st)))
                sys.exit(1)
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("profile run done: 3 x %s (variant %d) NN %d^3, then a 1 GiB copy" % (form, args.profile_variant, m))
        return

    results = {}
    flops = 8.0 * m * n * k
    for ops in args.ops.split(","):
        for name in args.variants.split(","):
            cfg = cgemm3m.config(name)
            C = C0.clone()
            st = call(cfg, ops, C)
            if st != 0:
                log("variant %s %s: status %s" % (name, ops, cgemm3m._lib.STATUS.get(st)))
                results["%s_%s" % (name, ops)] = {"status": cgemm3m._lib.STATUS.get(st, st)}
                continue
            for _ in range(args.warmup):
                call(cfg, ops, C)
            torch.cuda.synchronize()
            h.set_timing(True)
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            ev0.record(stream)
            for _ in range(args.steps):
                call(cfg, ops, C)
            ev1.record(stream)
            torch.cuda.synchronize()
            wall = (time.perf_counter() - t0) / args.steps * 1e3
            ev_ms = ev0.elapsed_time(ev1) / args.steps
            stages = h.timing()
            h.set_timing(False)
            calls = max(1, stages.pop("calls"))
            per_stage = {s: v / calls for s, v in stages.items() if v > 0}
            r = {"ms_per_step": wall, "ms_events": ev_ms, "tflops": flops / (wall * 1e-3) / 1e12,
                 "stage_ms": per_stage}
            gemm_ms = per_stage.get("gemm", 0.0)
            if gemm_ms > 0:
                ach = MFMA_FLOPS_PER_MNK[name] * m * n * k / (gemm_ms * 1e-3) / 1e12
                r["gemm_tflops"] = ach
                r["gemm_frac_of_peak"] = ach / PEAK[name]
            if not args.no_errors:
                Cr = C0.clone()
                call(cfg, ops, Cr)
                plan = h.last_plan(sync=True)
                res = cgemm3m.residual(h, ops[0], ops[1], m, n, k, alpha, P(A), m, P(B), m, beta, P(C0), P(Cr), m)
                Cr2 = C0.clone()
                call(cfg, ops, Cr2)
                torch.cuda.synchronize()
                r["identical_reruns"] = bool(torch.equal(Cr.view(torch.int32), Cr2.view(torch.int32)))
                r["err_vs_fp64"] = res
                r["plan"] = plan
                del Cr, Cr2
            results["%s_%s" % (name, ops)] = r
            log("%s %s: %.3f ms/step, %.1f complex TFLOP/s, stages %s, err %s" % (
                name, ops, wall, r["tflops"], {s: round(v, 3) for s, v in per_stage.items()},
                r.get("err_vs_fp64")))
            del C
            torch.cuda.empty_cache()

    # Row-scaled stress variant (rows of op(A) scaled by 10^u, u ~ U[-3, 3]): error only.
    if not args.no_errors and "NN" in args.ops.split(","):
        u = torch.rand(m, generator=gen, device=dev) * 6 - 3
        scale = torch.pow(10.0, u).to(torch.complex64)
        As = (A.view(k, m) * scale.view(1, m)).reshape(-1).contiguous()
        for name in args.variants.split(","):
            cfg = cgemm3m.config(name)
            Cr = C0.clone()
            if call(cfg, "NN", Cr, Aop=As) != 0:
                continue
            res = cgemm3m.residual(h, "N", "N", m, n, k, alpha, P(As), m, P(B), m, beta, P(C0), P(Cr), m)
            results.setdefault("row_scaled_err_vs_fp64", {})[name] = res
            log("row-scaled %s err %s" % (name, res))
        del As

    if not args.no_vendor and "NN" in args.ops.split(","):
        results["torch_matmul_complex64"] = vendor_bench(h, torch, cgemm3m, A, B, C0, m, n, k, args)
        log("torch.matmul complex64", results["torch_matmul_complex64"])

    if not args.no_autograd:
        results["autograd_2048x1024x4096"] = autograd_bench(torch, cgemm3m, dev, args)
        log("autograd", results["autograd_2048x1024x4096"])

    if not args.no_grouped:
        results["grouped_512"] = grouped_bench(h, torch, cgemm3m, dev)
        log("grouped", results["grouped_512"])

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(A, B, C0, m, n, k, alpha, beta, args.cpu_seconds)
        log("cpu baseline", cpu)

    main_r = results.get("fp32_NN", {})
    value = main_r.get("tflops")
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_gemm3m_f32_8192.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roof = None
    if "gemm_tflops" in main_r:
        roof = {"bound": "mfma", "achieved": round(main_r["gemm_tflops"], 2), "peak": PEAK["fp32"],
                "unit": "TFLOP/s", "frac": round(main_r["gemm_frac_of_peak"], 4), "traffic": traffic,
                "kernel": "gemm3m_f32_direct_kernel<128,128> (6mnk FP32 MFMA flops per launch)"}
    bf = results.get("bf16x3_NN", {})
    line = {
        "metric": METRIC,
        "value": round(value, 3) if value else None,
        "unit": "complex TFLOP/s (8mnk)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_r.get("ms_per_step", 0.0), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: re/im i.i.d. uniform[-1,1] (seeded torch generator on device)",
        "config": {"workload": "aclblasCgemm3m m=n=k=%d op=(N,N) alpha=(1,0.5) beta=(0.5,-0.25) "
                               "column-major complex64, lda=ldb=ldc=%d" % (m, m),
                   "m": m, "n": n, "k": k},
        "roofline": roof,
        "cpu_baseline": cpu,
        "bf16x3_over_fp32": (round(bf["tflops"] / value, 3) if bf.get("tflops") and value else None),
        "variants": results,
    }
    print(json.dumps(line))


def vendor_bench(h, torch, cgemm3m, A, B, C0, m, n, k, args):
    """The same product through torch.matmul on complex64 (the ROCm BLAS library's complex GEMM,
    four real products, no alpha/beta): C = A B in column-major storage is C^T = B^T A^T on the
    row-major views, so no operand is transposed or copied. A reported comparison on the same
    GPU, not a baseline of the metric."""
    P = cgemm3m.torch_ptr
    Bt, At = B.view(n, k), A.view(k, m)
    out = torch.empty(n, m, dtype=torch.complex64, device=A.device)
    for _ in range(2):
        torch.matmul(Bt, At, out=out)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    steps = max(3, args.steps // 2)
    ev0.record()
    for _ in range(steps):
        torch.matmul(Bt, At, out=out)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    r = {"ms_per_step": ms, "tflops": 8.0 * m * n * k / (ms * 1e-3) / 1e12,
         "op": "torch.matmul complex64 (N,N), alpha=1, beta=0"}
    if not args.no_errors:
        r["err_vs_fp64"] = cgemm3m.residual(h, "N", "N", m, n, k, 1.0, P(A), m, P(B), k, 0.0, P(C0), P(out), m)
    return r


def autograd_bench(torch, cgemm3m, dev, args, m=2048, k=1024, n=4096, seed=777):
    """The autograd configuration: Y = alpha A B + beta C, A m x k, B k x n, C m x n complex64
    (re/im uniform[-1,1]), alpha = (0.8, 0.6), beta = (0.5, 0); one step = the forward and the
    backward (dA, dB, dC) through the Cgemm3m kernels (FP32 and bf16x3), and the same step
    through torch.matmul autograd. Work per step: three complex products, 3 x 8mnk."""
    from cgemm3m.autograd import cgemm3m_mm

    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)

    def u(*shape):
        return (torch.rand(*shape, 2, generator=gen, device=dev) * 2 - 1).view(torch.complex64).reshape(shape)

    A, B, C, gY = u(m, k), u(k, n), u(m, n), u(m, n)
    alpha, beta = complex(0.8, 0.6), complex(0.5, 0.0)
    out = {}
    forms = {
        "fp32": lambda Ar, Br, Cr: cgemm3m_mm(Ar, Br, Cr, alpha, beta, "fp32"),
        "bf16x3": lambda Ar, Br, Cr: cgemm3m_mm(Ar, Br, Cr, alpha, beta, "bf16x3"),
        "torch_matmul": lambda Ar, Br, Cr: alpha * (Ar @ Br) + beta * Cr,
    }
    steps = max(3, args.steps // 2)
    for name, f in forms.items():
        Ar, Br, Cr = (x.clone().requires_grad_(True) for x in (A, B, C))

        def step():
            for x in (Ar, Br, Cr):
                x.grad = None
            f(Ar, Br, Cr).backward(gY)

        for _ in range(2):
            step()
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(steps):
            step()
        ev1.record()
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / steps
        out[name] = {"ms_per_step": ms, "tflops": 3 * 8.0 * m * n * k / (ms * 1e-3) / 1e12}
    out["step"] = "forward + backward of alpha*A@B + beta*C, A %dx%d, B %dx%d" % (m, k, k, n)
    return out


def grouped_problems(torch, cgemm3m, dev, count=512, seed=4242):
    """512 uneven problems (m, n in 16*{2..32}, k in 64*{1..32}, ops uniform over {N,T,C}^2,
    alpha = 1, beta = 0), seeded; returns (problems, buffers to keep alive, complex flops)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    probs, keep = [], []
    flops = 0.0
    for _ in range(count):
        ta, tb = str(rng.choice(list("NTC"))), str(rng.choice(list("NTC")))
        m = int(rng.integers(2, 33)) * 16
        n = int(rng.integers(2, 33)) * 16
        k = int(rng.integers(1, 33)) * 64
        lda, ldb = (m if ta == "N" else k), (k if tb == "N" else n)
        A = (torch.rand(2 * m * k, generator=gen, device=dev) * 2 - 1).view(torch.complex64)
        B = (torch.rand(2 * k * n, generator=gen, device=dev) * 2 - 1).view(torch.complex64)
        C = torch.zeros(m * n, dtype=torch.complex64, device=dev)
        keep += [A, B, C]
        probs.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=1.0, A=cgemm3m.torch_ptr(A), lda=lda,
                          B=cgemm3m.torch_ptr(B), ldb=ldb, beta=0.0, C=cgemm3m.torch_ptr(C), ldc=m))
        flops += 8.0 * m * n * k
    return probs, keep, flops


def grouped_bench(h, torch, cgemm3m, dev, count=512, seed=4242):
    """The grouped_problems batch: one grouped call, the same call replayed from a hipGraph,
    and a loop of one aclblasCgemm3m call per problem, both as the plain call (which, like the
    grouped members, turns the rounding-correction screen on for k >= 1024) and with the
    correction off. Each is timed in steady state: warm-up calls first (the GPU's clock ramps
    up over the first tens of milliseconds of load), then the timed ones. Complex TFLOP/s
    counts sum(8 m n k)."""
    probs, keep, flops = grouped_problems(torch, cgemm3m, dev, count, seed)
    stream = torch.cuda.current_stream()

    def timed(fn, warm, reps):
        for _ in range(warm):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    def loop(cfg=None):
        for p in probs:
            cgemm3m.cgemm3m(h, p["ta"], p["tb"], p["m"], p["n"], p["k"], 1.0, p["A"], p["lda"], p["B"], p["ldb"],
                            0.0, p["C"], p["ldc"], cfg)

    t_group = timed(lambda: cgemm3m.grouped(h, probs), 30, 20)
    t_loop = timed(loop, 3, 5)
    no_corr = cgemm3m.config(correction=-1)
    t_loop_nc = timed(lambda: loop(no_corr), 3, 5)
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        h.set_stream(side)
        cgemm3m.grouped(h, probs)
    torch.cuda.synchronize()
    with torch.cuda.graph(graph):
        h.set_stream(torch.cuda.current_stream())
        cgemm3m.grouped(h, probs)
    h.set_stream(stream)
    t_graph = timed(graph.replay, 40, 30)
    return {"problems": count, "complex_tflop": flops / 1e12, "ms_grouped": t_group, "ms_graph_replay": t_graph,
            "ms_per_problem_loop": t_loop, "ms_per_problem_loop_no_correction": t_loop_nc,
            "tflops_grouped": flops / (t_group * 1e-3) / 1e12,
            "tflops_graph": flops / (t_graph * 1e-3) / 1e12, "tflops_loop": flops / (t_loop * 1e-3) / 1e12,
            "speedup_graph_vs_loop": t_loop / t_graph, "speedup_graph_vs_loop_no_correction": t_loop_nc / t_graph}


def cpu_baseline(A, B, C0, m, n, k, alpha, beta, seconds):
    """Netlib-order oracle on the host cores, on the first `cols` output columns of the same
    inputs (columns are independent in Netlib's order, so this is a bounded sample)."""
    import numpy as np

    import oracle

    threads = min(16, len(os.sched_getaffinity(0)))
    Ah = A.cpu().numpy()
    cols = 2
    rate = None
    elapsed = 0.0
    while True:
        Bh = B[: cols * k].cpu().numpy()
        Ch = C0[: cols * m].cpu().numpy()
        t0 = time.perf_counter()
        oracle.cgemm("N", "N", m, cols, k, alpha, Ah, m, Bh, k, beta, Ch, m, threads=threads)
        elapsed = time.perf_counter() - t0
        rate = 8.0 * m * cols * k / elapsed
        if elapsed * 4 > seconds or cols >= n:
            break
        cols = min(n, max(cols * 2, int(cols * seconds / max(elapsed, 1e-3) * 0.9)))
    # "cores" is the threads used (the contract's meaning); the host's own count beside it
    return {"value": round(rate / 1e12, 6), "unit": "complex TFLOP/s (8mnk)", "cores": threads, "threads": threads,
            "host_cores": os.cpu_count(), "kind": "port",
            "sample": "Netlib-order cgemm restatement (oracle/oracle.c, -O2 -ffp-contract=off, OpenMP over "
                      "columns, %d threads on a host with %d cores) on the first %d of %d output columns of the "
                      "same %d^3 NN inputs, %.1f s; the rate is 8*m*cols*k over that time (Netlib's columns are "
                      "independent, so the rate of a column sample is the whole product's)"
                      % (threads, os.cpu_count(), cols, n, m, elapsed)}


if __name__ == "__main__":
    main()
