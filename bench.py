#!/usr/bin/env python
"""TVAL3 reconstructions per second on the 2-D ultrasound ring, one MI355X.

Workload: 256 transducers on the circle inscribed in a 256x256 grid, all 32,640
emitter-receiver pairs (i < j), on-the-fly Bresenham projector with weight 1 per
traversed pixel, disk phantom with seeded inclusions, y = Ax + 0.5 % noise,
float32, maxit 200, default TVAL3 options (mu = 2^8, beta = 2^5).

A "step" is one batched solve: B independent reconstructions (one measurement
vector each, shared geometry) advanced together by the engine's 64-lane kernels.
value = B * steps / wall time of the timed steps (inputs already in HBM).  The
plain single-instance FP32 path (W = 1 kernels) is timed beside it, and the CPU
oracle (the C++/OpenMP restatement of the reference) on the host cores.

Only one GPU is used; --gpus above 1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "tval3-cpu_amd"), os.path.join(ROOT, "oracle")]

METRIC = ("TVAL3 recons/s, 256² ring (plain FP32 vs batched BF16, ≤1e-3 rel err); "
          "A/Aᵀ GB/s")
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def csr_equiv_bytes(nnz, m, n):
    """Algorithmic bytes of one A or A^T product, CSR-equivalent (4-byte value +
    4-byte index per entry, row pointers, x and y)."""
    return 8 * nnz + 4 * (m + 1) + 4 * n + 4 * m


def cpu_baseline(geo, b, opts, budget_s):
    import numpy as np  # noqa: F401
    import tv3_oracle as O
    threads = len(os.sched_getaffinity(0))
    O.set_threads(threads)
    op = O.Op.geometry(geo)
    dims = (geo.rv[0], geo.rv[1], 1)
    runs, t0 = 0, time.perf_counter()
    while True:
        O.solve(op, b, dims, opts)
        runs += 1
        el = time.perf_counter() - t0
        if el >= budget_s or runs >= 20:
            break
    return {"value": runs / el, "unit": "recons/s", "cores": O.get_threads(),
            "kind": "port",
            "sample": f"{runs} full 200-iteration reconstructions of the 256x256 "
                      f"ring workload (C++/OpenMP oracle, {el:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--plain-runs", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--maxit", type=int, default=200)
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run a device-to-device copy of this many MiB "
                         "(calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the "
              "multi-GPU sharding of this path is design-only.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import tval3_amd
    from tval3_amd import synth

    ctx = tval3_amd.Context(0)
    G = args.grid
    geo = synth.ring_geometry(G, G)
    t0 = time.perf_counter()
    plan = ctx.plan_geometry(geo)
    log(f"plan: m={plan.m} n={plan.n} nnz={plan.nnz} ({time.perf_counter() - t0:.2f} s)")
    x = synth.disk_phantom(G, seed=0)
    y = ctx.gemv(plan, x).cpu().numpy()
    B = args.batch
    bs_host = [synth.add_noise(y, 0.005, seed=i) for i in range(B)]
    bs = [ctx.tensor(b) for b in bs_host]
    dims = (G, G, 1)
    opts = tval3_amd.default_options(maxit=args.maxit)
    torch.cuda.synchronize()

    # ---- plain FP32 path: one reconstruction at a time (W = 1 kernels)
    U0, info0, _ = ctx.solve(plan, bs[0], dims, opts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.plain_runs):
        ctx.solve(plan, bs[i % B], dims, opts)
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / args.plain_runs
    log(f"plain: {1e3 * t_plain:.1f} ms/recon, iters={info0.total_iters}")

    # ---- batched path: B reconstructions per step (64-lane kernels)
    plans = [plan] * B
    dl = [dims] * B
    for _ in range(args.warmup):
        Us, infos = ctx.solve_batch(plans, bs, dl, opts)
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        Us, infos = ctx.solve_batch(plans, bs, dl, opts)
        log(f"step {k + 1}/{args.steps} done")
    torch.cuda.synchronize()
    t_steps = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile(False)
    value = B * args.steps / t_steps
    rel_b0 = float(torch.linalg.norm(Us[0] - U0) / torch.linalg.norm(U0))
    iters = [inf.total_iters for inf in infos]

    # ---- roofline of the projector (A and A^T, CSR-equivalent bytes)
    Bpad = ((B + 63) // 64) * 64
    per_inst = csr_equiv_bytes(plan.nnz, plan.m, plan.n)
    ops = {}
    for k, (ms, cnt) in prof.items():
        avg = ms / max(cnt, 1)
        ops[k] = {"launches": cnt, "ms_per_launch": avg,
                  "GBps": Bpad * per_inst / (avg * 1e-3) / 1e9 if cnt else None}
    dom = max(ops, key=lambda k: prof[k][0])
    ach = ops[dom]["GBps"]
    roofline = {"bound": "hbm", "kernel": "k_fwd_w64" if dom == "A" else "k_bwd_w64",
                "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBPS if ach else None, "traffic": None,
                "bytes_per_launch": Bpad * per_inst}

    if args.copy_mb:
        src = torch.ones(args.copy_mb * 262144, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(geo, bs_host[0], opts, args.cpu_seconds)
        log(f"cpu: {cpu['value']:.3f} recons/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC, "value": value, "unit": "recons/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t_steps / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded disk phantom with inclusions, 0.5% noise",
        "config": {"workload": f"2D USCT ring {G}x{G}, {plan.m} rays, "
                               f"on-the-fly Bresenham, maxit {args.maxit}",
                   "batch": B, "grid": G, "rays": plan.m, "nnz": plan.nnz,
                   "maxit": args.maxit},
        "plain_fp32_recons_per_s": 1.0 / t_plain,
        "batched_over_plain": value * t_plain,
        "batch_vs_plain_rel_l2": rel_b0,
        "iters_min_max": [min(iters), max(iters)],
        "operators": ops,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
