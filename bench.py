#!/usr/bin/env python
"""TVAL3 reconstructions per second on one MI355X (default: the 2-D ultrasound
ring, 256x256).

Workloads (--workload):
  ring         256 transducers on the circle inscribed in a 256x256 grid, all
               32,640 emitter-receiver pairs (i < j), on-the-fly Bresenham
               projector (weight 1 per traversed pixel), disk phantom with
               seeded inclusions, y = Ax + 0.5 % noise, float32, maxit 200,
               default TVAL3 options (mu = 2^8, beta = 2^5).  A step is one
               batched solve of B reconstructions sharing the geometry (64-lane
               kernels, B = 448 by default); the plain single-reconstruction
               path (W = 1) is timed beside it.  [default; the headline metric]
  dense-batch  64 independent dense reconstructions of uneven size (grids from
               {64..256}^2, m ~ U[20 %, 40 %] n), A ~ N(0, 1/m) stored bf16, one
               batched solve; per-instance error against the FP64 solve of the
               same (bf16-rounded) matrices on a sample of instances.
  hemisphere   1024 Fibonacci transducers on the lower hemisphere of a 128^3
               volume, 523,776 rays, 3-D anisotropic TV, maxit 100, one solve.
  dense        64x64 image, m = 1229, dense FP32 A, isotropic TV, maxit 300.
  dense-layers the reference's 2-D multi-layer mode: --layers (64) images of
               64x64 stacked along y sharing one A (m = 1229 per layer),
               isotropic TV, maxit 300; both products are FP64-MFMA GEMMs
               over the layers; value = layered solves/s (images/s = layers x).
  ring-f64     the 256x256 ring with the double/double solver (FP64 vectors
               and scalars, tv3_solve_f64), one reconstruction at a time,
               maxit 200; CPU baseline: the oracle's double/double solve.
  diff         gradients of |U_K - x|^2 in b and mu through K = 20 unrolled
               iterations (checkpoints every 5), 128-transducer ring on 128x128.

value = reconstructions per second of the timed steps (inputs already in HBM).
The CPU oracle (the C++/OpenMP restatement of the reference) is timed on the
host cores on a bounded sample of the same workload.  The roofline object
reports the dominant measurement-operator kernel: its compulsory bytes per
launch over its mean launch time, taken live with HIP events on the solver's
stream.  Compulsory bytes: the operator once for every group of instances
that computed (dense: the stored A; CSR / CSC: indices, values and row
pointers; the on-the-fly projector: its 32-byte ray descriptors for A x, its
voxel-major ray lists for A^T) plus the input and output vectors of every
instance that computed.  The per-instance CSR-equivalent rate (8 bytes per
entry for every instance, as if each instance streamed its own CSR matrix)
and, for the projector, the rate of the 4-byte gathers per entry and
instance (served by the caches) are reported beside it, labelled.  "traffic"
is the HBM bytes per launch measured with rocprofv3 (FETCH_SIZE, WRITE_SIZE,
calibrated on a copy kernel) when a committed profile summary for the
workload exists.

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
FP64_MFMA_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# the single-instance projector transpose: the staged kernel unless
# TV3_STAGE=0 selects the unstaged one
W1_TRANSPOSE = "k_bwd_w1" if os.environ.get("TV3_STAGE", "1") == "0" else "k_bwd_w1s"


def traffic_from_profiles(workload, kernel):
    """HBM bytes per launch of `kernel` from profiles/traffic_<workload>.json
    (written from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_time(fn, budget_s, max_runs=20):
    import tv3_oracle as O
    O.set_threads(len(os.sched_getaffinity(0)))
    runs, t0 = 0, time.perf_counter()
    while True:
        fn()
        runs += 1
        el = time.perf_counter() - t0
        if el >= budget_s or runs >= max_runs:
            return runs, el, O.get_threads()


def roofline(prof, kernels, workload, plan=None, t_steps=None, steps=None):
    """Per-operator figures from the library's launch profile.  Each launch
    is charged its compulsory bytes (the operator once per computing group,
    the vectors of each computing instance), so achieved = bytes / time is a
    physical rate.  With a plan, the per-instance CSR-equivalent rate and,
    for the projector, the cache-served gather rate are added (labelled).
    With t_steps, the compulsory bytes of all products per step over the
    step's wall time are checked against the HBM peak."""
    ops = {}
    for k, (ms, cnt, by, lanes) in prof.items():
        ops[k] = {"launches": cnt, "ms_per_launch": ms / max(cnt, 1),
                  "bytes_per_launch": by / max(cnt, 1),
                  "instances_per_launch": lanes / max(cnt, 1),
                  "GBps": by / (ms * 1e-3) / 1e9 if ms > 0 else None}
        if plan is not None and ms > 0 and workload in ("ring", "hemisphere", "diff", "ring-f64"):
            rows = plan.m if k == "A" else plan.n
            csr = 8 * plan.nnz + 4 * (rows + 1) + 4 * (plan.m + plan.n)
            ops[k]["csr_equivalent_GBps_per_instance_operator"] = lanes * csr / (ms * 1e-3) / 1e9
            es = 8 if workload == "ring-f64" else 4
            ops[k]["cache_gather_GBps"] = lanes * es * plan.nnz / (ms * 1e-3) / 1e9
    dom = max(ops, key=lambda k: prof[k][0])
    ach = ops[dom]["GBps"]
    kern = kernels[dom]
    traffic = traffic_from_profiles(workload, kern)
    roof = {"bound": "hbm", "kernel": kern, "achieved": ach,
            "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBPS if ach else None,
            "traffic": traffic,
            "bytes_per_launch": ops[dom]["bytes_per_launch"],
            "ms_per_launch": ops[dom]["ms_per_launch"],
            "traffic_over_bytes": (traffic / ops[dom]["bytes_per_launch"]
                                   if traffic and ops[dom]["bytes_per_launch"] else None)}
    if t_steps:
        tot = sum(prof[k][2] for k in prof) / steps
        rate = tot / (t_steps / steps) / 1e9
        roof["products_GBps_over_wall"] = rate
        if rate > HBM_PEAK_GBPS:
            raise RuntimeError(f"compulsory product bytes per step over the step's wall time "
                               f"({rate:.0f} GB/s) exceed the HBM peak: byte accounting is wrong")
    return ops, roof


# ---------------------------------------------------------------------------
def run_ring(args, ctx):
    import torch
    import tval3_amd
    from tval3_amd import synth
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
    opts = tval3_amd.default_options(maxit=args.maxit, graph=args.graph)
    torch.cuda.synchronize()

    # plain FP32 path: one reconstruction at a time (W = 1 kernels)
    U0, info0, _ = ctx.solve(plan, bs[0], dims, opts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.plain_runs):
        ctx.solve(plan, bs[i % B], dims, opts)
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / args.plain_runs
    log(f"plain: {1e3 * t_plain:.1f} ms/recon, iters={info0.total_iters}")

    # batched path: B reconstructions per step (64-lane kernels)
    plans, dl = [plan] * B, [dims] * B
    for _ in range(args.warmup):
        Us, infos = ctx.solve_batch(plans, bs, dl, opts)
    # one untimed step counts the products' compulsory bytes and lanes on the
    # device; the timed steps (the same deterministic launches) time the
    # products with events only, so no counting kernel runs in them
    ctx.profile(1)
    ctx.solve_batch(plans, bs, dl, opts)
    counted = ctx.profile_read()
    torch.cuda.synchronize()
    ctx.profile(3)
    t0 = time.perf_counter()
    for k in range(args.steps):
        Us, infos = ctx.solve_batch(plans, bs, dl, opts)
        log(f"step {k + 1}/{args.steps} done")
    torch.cuda.synchronize()
    t_steps = time.perf_counter() - t0
    timed = ctx.profile_read()
    ctx.profile(False)
    prof = {}
    for k in timed:
        ms, cnt = timed[k][0], timed[k][1]
        _, cnt1, by1, ln1 = counted[k]
        if cnt != cnt1 * args.steps:
            raise RuntimeError(f"{k}: {cnt} timed launches, {cnt1} per counted step")
        prof[k] = (ms, cnt, by1 * args.steps, ln1 * args.steps)
    args.one_step = lambda: ctx.solve_batch(plans, bs, dl, opts)
    value = B * args.steps / t_steps
    rel_b0 = float(torch.linalg.norm(Us[0] - U0) / torch.linalg.norm(U0))
    iters = [inf.total_iters for inf in infos]
    ops, roof = roofline(prof, {"A": "k_fwd_w64q", "At": "k_bwd_w64q"}, "ring", plan,
                         t_steps, args.steps)
    # the walk's gathers (one 256-byte row of 64 instances per entry) are
    # served by L2 and the last-level cache
    roof["cache_gather_GBps"] = ops["A"]["cache_gather_GBps"]
    roof["voxel_visits_per_s"] = (ops["A"]["instances_per_launch"] * plan.nnz
                                  / (ops["A"]["ms_per_launch"] * 1e-3))
    import numpy as np
    import tv3_oracle as O
    op = O.Op.geometry(geo)
    fp64 = None
    if not args.no_fp64:
        fp64 = ring_precision(args, ctx, plan, op, bs_host, Us, dims, opts)
        log(f"fp64: {json.dumps(fp64['summary'])}")
    cpu = None
    if not args.no_cpu:
        runs, el, thr = cpu_time(lambda: O.solve(op, bs_host[0], dims, opts), args.cpu_seconds)
        cpu = {"value": runs / el, "unit": "recons/s", "cores": thr, "kind": "port",
               "sample": f"{runs} full {args.maxit}-iteration reconstructions of the "
                         f"{G}x{G} ring workload (C++/OpenMP oracle, {el:.1f} s)"}
        log(f"cpu: {cpu['value']:.3f} recons/s on {thr} threads")
    out = {
        "value": value, "unit": "recons/s", "steps": args.steps,
        "ms_per_step": 1e3 * t_steps / args.steps, "dtype": "f32",
        "data": "synthetic: seeded disk phantom with inclusions, 0.5% noise",
        "config": {"workload": f"2D USCT ring {G}x{G}, {plan.m} rays, "
                               f"on-the-fly Bresenham, maxit {args.maxit}",
                   "batch": B, "grid": G, "rays": plan.m, "nnz": plan.nnz,
                   "maxit": args.maxit},
        "plain_fp32_recons_per_s": 1.0 / t_plain,
        "batched_over_plain": value * t_plain,
        "batch_vs_plain_rel_l2": rel_b0,
        "iters_min_max": [min(iters), max(iters)],
        "measured": ("batched FP32 reconstructions (FP32 vectors, T_scal float) with the "
                     "on-the-fly weight-1 projector; bf16 storage applies to dense A "
                     "(--workload dense-batch); error vs FP64 over the sample in "
                     "precision_vs_fp64"),
        "precision_vs_fp64": fp64,
        "operators": ops, "roofline": roof, "cpu_baseline": cpu,
    }
    return out


def ring_precision(args, ctx, plan, op, bs_host, Us, dims, opts):
    """FP32 batched results against the FP64 solve (tv3_solve_batch_f64 on
    the GPU) over a sample of instances, and, for each, the first iteration
    whose line-search decision (backtracks, fallback) differs between the
    FP32 and FP64 trajectories, with the two iterates' distance one
    iteration before it.  Instance 0 is also checked against the CPU oracle
    (the FP32 reference arithmetic) bit for bit."""
    import numpy as np
    import torch
    K = min(args.fp64_sample_ring, len(bs_host))
    b64 = [ctx.tensor(bs_host[i], torch.float64) for i in range(K)]
    U64, _ = ctx.solve_batch([plan] * K, b64, [dims] * K, opts, f64=True)
    per = []
    for i in range(K):
        u32 = Us[i].cpu().numpy().astype(np.float64)
        u64 = U64[i].cpu().numpy()
        nrm = np.linalg.norm(u64)
        _, _, t32 = ctx.solve(plan, ctx.tensor(bs_host[i]), dims, opts, trace_cap=opts.maxit)
        _, _, t64 = ctx.solve(plan, b64[i], dims, opts, trace_cap=opts.maxit, f64=True)
        k = next((a.iter for a, c in zip(t32, t64)
                  if (a.backtracks, a.fallback) != (c.backtracks, c.fallback)), None)
        rec = {"instance": i, "rel_l2": float(np.linalg.norm(u32 - u64) / nrm),
               "fallbacks_fp32": sum(t.fallback for t in t32),
               "fallbacks_fp64": sum(t.fallback for t in t64),
               "first_decision_diff_iter": k}
        if k is not None and k > 1:
            o2 = tval3_amd_options(opts, maxit=k - 1)
            a32, _, _ = ctx.solve(plan, ctx.tensor(bs_host[i]), dims, o2)
            a64, _, _ = ctx.solve(plan, b64[i], dims, o2, f64=True)
            a64 = a64.cpu().numpy()
            rec["rel_l2_one_iter_before_diff"] = float(
                np.linalg.norm(a32.cpu().numpy().astype(np.float64) - a64) / np.linalg.norm(a64))
        per.append(rec)
    rels = np.array([r["rel_l2"] for r in per])
    before = [r["rel_l2_one_iter_before_diff"] for r in per if "rel_l2_one_iter_before_diff" in r]
    # the FP64 solve's own sensitivity: b perturbed by 1e-12 relative
    rng = np.random.default_rng(5)
    bp = ctx.tensor(bs_host[0].astype(np.float64) * (1 + 1e-12 * rng.standard_normal(plan.m)),
                    torch.float64)
    Up, _, _ = ctx.solve(plan, bp, dims, opts, f64=True)
    u640 = U64[0].cpu().numpy()
    U32o, _, _ = O_solve_oracle(op, bs_host[0], dims, opts)
    summary = {"instances": K,
               "fp32_vs_fp64_rel_l2_min_median_max": [float(rels.min()), float(np.median(rels)),
                                                      float(rels.max())],
               "instances_within_1e-3": int((rels <= 1e-3).sum()),
               "rel_l2_one_iter_before_first_decision_diff_max": (float(max(before))
                                                                  if before else None),
               "fp64_vs_fp64_b_perturbed_1e-12": float(np.linalg.norm(
                   Up.cpu().numpy() - u640) / np.linalg.norm(u640)),
               "batched_vs_oracle_fp32_instance0": float(np.linalg.norm(
                   Us[0].cpu().numpy() - U32o) / np.linalg.norm(U32o))}
    return {"summary": summary, "per_instance": per}


def tval3_amd_options(opts, **kw):
    import tval3_amd
    d = opts.as_dict()
    d.update(kw)
    return tval3_amd.default_options(**d)


def O_solve_oracle(op, b, dims, opts):
    import tv3_oracle as O
    return O.solve(op, b, dims, opts)


def run_dense_batch(args, ctx):
    import numpy as np
    import torch
    import tval3_amd
    from tval3_amd import synth
    B = args.batch if args.batch_set else 64
    sizes = synth.batched_dense_sizes(B, seed=0)
    gen = torch.Generator(device="cuda")
    plans, bs, dims, keep = [], [], [], []
    tot_elems = 0
    for k, (g, m) in enumerate(sizes):
        n = g * g
        gen.manual_seed(1000 + k)
        Ab = torch.empty((m, n), dtype=torch.bfloat16, device="cuda")
        step = max(1, (1 << 28) // n)
        for r0 in range(0, m, step):
            r1 = min(m, r0 + step)
            Ab[r0:r1] = (torch.randn((r1 - r0, n), generator=gen, device="cuda")
                         * (1.0 / np.sqrt(m))).to(torch.bfloat16)
        x = torch.from_numpy(synth.ellipse_phantom(g, g, 8, seed=k)).cuda()
        y = (Ab.float() @ x) if m * n < (1 << 31) else torch.cat(
            [Ab[r0:r0 + 4096].float() @ x for r0 in range(0, m, 4096)])
        b = torch.from_numpy(synth.add_noise(y.cpu().numpy(), 0.01, seed=k)).cuda()
        plans.append(ctx.plan_dense(Ab, bf16=True))
        bs.append(b); dims.append((g, g, 1)); keep.append((Ab, x))
        tot_elems += m * n
    torch.cuda.synchronize()
    log(f"dense-batch: {B} instances, {tot_elems / 1e9:.2f} G bf16 entries "
        f"({2 * tot_elems / 2**30:.1f} GiB)")
    # FP32 vectors, FP64 step-size / line-search scalars; --vectors f64: FP64
    # vectors and scalars on the same bf16 A (the FP64 solve of the rounded
    # matrices)
    opts = tval3_amd.default_options(maxit=args.maxit, scal_double=1, graph=args.graph)
    f64v = args.vectors == "f64"
    if f64v:
        bs = [b.to(torch.float64) for b in bs]
    for _ in range(args.warmup):
        ctx.solve_batch(plans, bs, dims, opts, f64=f64v)
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        Us, infos = ctx.solve_batch(plans, bs, dims, opts, f64=f64v)
        log(f"step {k + 1}/{args.steps} done")
    torch.cuda.synchronize()
    t_steps = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile(False)
    args.one_step = lambda: ctx.solve_batch(plans, bs, dims, opts, f64=f64v)
    value = B * args.steps / t_steps
    kern = ({"A": "k_dense_fwd_db", "At": "k_dense_bwd_part_db"} if f64v
            else {"A": "k_dense_fwd_w1", "At": "k_dense_bwd_blk"})
    ops, roof = roofline(prof, kern, "dense-batch" + ("-f64" if f64v else ""), None, t_steps,
                         args.steps)
    # the same batch replayed from a hipGraph (tv3_options.graph = 1: every
    # pair of iterations is one graph launch, the host reads the stop flags
    # once per launch); timed apart, since launches inside a graph carry no
    # events, and held to the host-driven result bit for bit
    graph = None
    if not args.graph:
        gopts = tval3_amd.default_options(maxit=args.maxit, scal_double=1, graph=1)
        ctx.solve_batch(plans, bs, dims, gopts, f64=f64v)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Ug, _ = ctx.solve_batch(plans, bs, dims, gopts, f64=f64v)
        torch.cuda.synchronize()
        tg = time.perf_counter() - t0
        same = all(torch.equal(a, b) for a, b in zip(Ug, Us))
        graph = {"recons_per_s": B / tg, "ms_per_step": 1e3 * tg,
                 "bit_identical_to_host_driven": bool(same)}
        log(f"graph mode: {B / tg:.2f} recons/s, bit-identical: {same}")
        if not same:
            raise RuntimeError("graph-mode batch differs from the host-driven batch")
        del Ug
    # every instance against the FP64 solve of the same bf16-rounded
    # matrices (tv3_solve_batch_f64 on the GPU, in chunks of FP64 copies of
    # A), and the smallest ones against the CPU oracle's FP32 solve (the
    # reference arithmetic, bit for bit)
    import tv3_oracle as O
    errs = {}
    if not args.no_fp64:
        chunk, used = [], 0.0
        for k in list(range(B)) + [None]:
            if k is not None:
                g, m = sizes[k]
                need = 8.0 * m * g * g / 2 ** 30
            if chunk and (k is None or used + need > args.fp64_chunk_gib):
                p64 = [ctx.plan_dense(keep[j][0].to(torch.float64), f64=True) for j in chunk]
                b64 = [bs[j].to(torch.float64) for j in chunk]
                # the reference: FP64 copies of the bf16-rounded A
                U64, _ = ctx.solve_batch(p64, b64, [dims[j] for j in chunk], opts, f64=True)
                for j, u in zip(chunk, U64):
                    u = u.cpu().numpy()
                    errs[j] = {"grid": sizes[j][0], "m": sizes[j][1], "gpu_vs_fp64": float(
                        np.linalg.norm(Us[j].cpu().numpy().astype(np.float64) - u) / np.linalg.norm(u))}
                del p64, b64, U64
                torch.cuda.empty_cache()
                chunk, used = [], 0.0
            if k is not None:
                chunk.append(k)
                used += need
    order = sorted(range(B), key=lambda k: sizes[k][0] ** 2 * sizes[k][1])
    for k in (order[:args.f64_sample] if not f64v else []):
        op = O.Op.dense(keep[k][0].float().cpu().numpy())
        U32, _, _ = O.solve(op, bs[k].cpu().numpy(), dims[k], opts)
        errs.setdefault(k, {"grid": sizes[k][0], "m": sizes[k][1]})["gpu_vs_oracle_fp32"] = float(
            np.linalg.norm(Us[k].cpu().numpy() - U32) / np.linalg.norm(U32))
    e64 = np.array([v["gpu_vs_fp64"] for v in errs.values() if "gpu_vs_fp64" in v])
    prec = {"instances_vs_fp64": int(e64.size),
            "rel_l2_min_median_max": ([float(e64.min()), float(np.median(e64)), float(e64.max())]
                                      if e64.size else None),
            "instances_within_1e-3": int((e64 <= 1e-3).sum()),
            "instances_within_1e-2": int((e64 <= 1e-2).sum()),
            "per_instance": errs}
    log(f"rel L2 vs FP64 over {e64.size} instances: {prec['rel_l2_min_median_max']}")
    cpu = None
    if not args.no_cpu:
        k = order[0]
        Ab, x = keep[k]
        op = O.Op.dense(Ab.float().cpu().numpy())
        bk = bs[k].cpu().numpy()
        runs, el, thr = cpu_time(lambda: O.solve(op, bk, dims[k], opts), args.cpu_seconds)
        # scale by the batch's entry count (the solve is A-bound on the CPU)
        m0, g0 = sizes[k][1], sizes[k][0]
        frac = (m0 * g0 * g0) / tot_elems * B
        cpu = {"value": runs / el * frac, "unit": "recons/s (batch-average size)",
               "cores": thr, "kind": "port",
               "sample": f"{runs} reconstructions of instance {k} ({g0}^2, m={m0}) "
                         f"in {el:.1f} s, scaled by entry count to the batch's "
                         f"average instance"}
    return {
        "value": value, "unit": "recons/s", "steps": args.steps,
        "ms_per_step": 1e3 * t_steps / args.steps,
        "dtype": "bf16 A, f64 vectors" if f64v else "bf16 A, f32 vectors",
        "data": "synthetic: A ~ N(0,1/m) rounded to bf16, 8-ellipse phantoms, 1% noise",
        "config": {"workload": f"{B} uneven dense 2D reconstructions (bf16 A), maxit {args.maxit}",
                   "batch": B, "bf16_entries": tot_elems, "maxit": args.maxit},
        "precision_vs_fp64": prec, "iters": [i.total_iters for i in infos],
        "graph_replay": graph,
        "operators": ops, "roofline": roof, "cpu_baseline": cpu,
    }


def run_diff(args, ctx):
    """Gradient of |U_K - x|^2 with respect to b and mu through K = 20 unrolled
    iterations (checkpoints every 5) on the 128-transducer ring, 128x128."""
    import numpy as np
    import torch
    from tval3_amd import diff, synth
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    G, Kt, K, ck = 128, 128, 20, 5
    geo = synth.ring_geometry(Kt, G)
    plan = ctx.plan_geometry(geo)
    x = synth.disk_phantom(G, seed=0)
    y = ctx.gemv(plan, x).cpu().numpy()
    b = ctx.tensor(synth.add_noise(y, 0.005, seed=0))
    xt = ctx.tensor(x)
    mu = 256.0
    alpha, first = diff.record(ctx, plan, b, (G, G), mu, K)

    def grad_once():
        bt = b.clone().requires_grad_(True)
        mt = torch.tensor(mu, dtype=torch.float64, device=b.device, requires_grad=True)
        U = diff.unrolled(ctx, plan, bt, mt, (G, G), K=K, ckpt=ck, alpha=alpha, first=first)
        L = ((U - xt) ** 2).sum()
        L.backward()
        return bt.grad, mt.grad

    for _ in range(args.warmup):
        grad_once()
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        grad_once()
    torch.cuda.synchronize()
    t_steps = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile(False)
    args.one_step = grad_once
    ops, roof = roofline(prof, {"A": "k_fwd_w1", "At": W1_TRANSPOSE}, "diff", plan,
                         t_steps, args.steps)
    roof["cache_resident"] = True
    cpu = None
    if not args.no_cpu:
        import diff_ref as R
        import tv3_oracle as O
        op = O.Op.geometry(geo)
        ptr, lin = op.walk()
        A, AT = R.sparse_from_walk(ptr, lin, geo.m, geo.n)
        bh = b.cpu().numpy().astype(np.float64)
        torch.set_num_threads(len(os.sched_getaffinity(0)))

        def cpu_grad():
            bt = torch.tensor(bh, requires_grad=True)
            mt = torch.tensor(mu, dtype=torch.float64, requires_grad=True)
            U = R.unrolled(A, AT, bt, mt, 32.0, G, G, K, alpha, first)
            ((U - torch.from_numpy(x.astype(np.float64))) ** 2).sum().backward()

        runs, t1 = 0, time.perf_counter()
        while time.perf_counter() - t1 < args.cpu_seconds and runs < 50:
            cpu_grad()
            runs += 1
        el = time.perf_counter() - t1
        cpu = {"value": runs / el, "unit": "gradients/s", "cores": torch.get_num_threads(),
               "kind": "port",
               "sample": f"{runs} FP64 torch-autograd gradients of the same unrolled map "
                         f"on the CPU ({el:.1f} s)"}
    return {
        "value": args.steps / t_steps, "unit": "gradients/s", "steps": args.steps,
        "ms_per_step": 1e3 * t_steps / args.steps, "dtype": "f32",
        "data": "synthetic: disk phantom, 0.5% noise",
        "config": {"workload": f"d|U_K - x|^2 / d(b, mu), K={K} unrolled iterations, "
                               f"checkpoints every {ck}, ring {Kt} transducers on {G}x{G}",
                   "rays": plan.m, "nnz": plan.nnz, "K": K, "ckpt": ck},
        "operators": ops, "roofline": roof, "cpu_baseline": cpu,
    }


def run_single(args, ctx, workload):
    import numpy as np
    import torch
    import tval3_amd
    from tval3_amd import synth
    import tv3_oracle as O
    if workload == "hemisphere":
        grid = args.grid if args.grid_set else 128
        geo = synth.hemisphere_geometry(1024, grid)
        plan = ctx.plan_geometry(geo)
        x = synth.ellipsoid_phantom(grid, 20, seed=0)
        y = ctx.gemv(plan, x).cpu().numpy()
        b = synth.add_noise(y, 0.005, seed=0)
        dims = (grid, grid, grid)
        maxit = args.maxit if args.maxit_set else 100
        opts = tval3_amd.default_options(maxit=maxit, graph=args.graph)
        kern = {"A": "k_fwd_w1", "At": W1_TRANSPOSE}
        op_fn = lambda: O.Op.geometry(geo)  # noqa: E731
        desc = f"3D USCT hemisphere {grid}^3, {plan.m} rays, maxit {maxit}"
    elif workload == "ring-f64":
        grid = args.grid if args.grid_set else 256
        geo = synth.ring_geometry(256, grid)
        plan = ctx.plan_geometry(geo)
        x = synth.disk_phantom(grid, seed=0)
        y = ctx.gemv(plan, x).cpu().numpy()
        b = synth.add_noise(y, 0.005, seed=0).astype(np.float64)
        dims = (grid, grid, 1)
        maxit = args.maxit if args.maxit_set else 200
        opts = tval3_amd.default_options(maxit=maxit, graph=args.graph)
        kern = {"A": "k_fwd_w1d", "At": "k_bwd_w1d"}
        op_fn = lambda: O.Op.geometry(geo)  # noqa: E731
        desc = f"2D USCT ring 256 transducers on {grid}x{grid}, FP64 vectors, maxit {maxit}"
    elif workload == "dense-layers":
        q = p = 64
        m, L = 1229, args.layers
        A = synth.gaussian_A(m, q * p, 0)
        x = np.concatenate([synth.ellipse_phantom(p, q, 8, l) for l in range(L)])
        b = np.concatenate([synth.add_noise(A @ x[l * q * p:(l + 1) * q * p], 0.01, l)
                            for l in range(L)])
        plan = ctx.plan_dense(A)
        dims = (q, L * p, 1)
        maxit = args.maxit if args.maxit_set else 300
        opts = tval3_amd.default_options(maxit=maxit, tv_iso=1, graph=args.graph)
        # the register-blocked kernels (one-tile k_dense_*_mfma with
        # TV3_MFMA_BLOCK=1,1)
        blk = os.environ.get("TV3_MFMA_BLOCK", "3,2").split(",")
        kern = {"A": "k_dense_fwd_mfma" + ("" if blk[0] == "1" else "b"),
                "At": "k_dense_bwd_mfma" + ("" if len(blk) > 1 and blk[1] == "1" else "b")}
        op_fn = lambda: O.Op.dense(A)  # noqa: E731
        desc = (f"2D dense multi-layer: {L} layers of 64x64 sharing one A (m={m} per layer), "
                f"isotropic TV, maxit {maxit}")
    else:
        q = p = 64
        m = 1229
        A = synth.gaussian_A(m, q * p, 0)
        x = synth.ellipse_phantom(p, q, 8, 0)
        b = synth.add_noise(A @ x, 0.01, 0)
        plan = ctx.plan_dense(A)
        dims = (q, p, 1)
        maxit = args.maxit if args.maxit_set else 300
        opts = tval3_amd.default_options(maxit=maxit, tv_iso=1, graph=args.graph)
        kern = {"A": "k_dense_fwd_w1", "At": "k_dense_bwd_bs"}
        op_fn = lambda: O.Op.dense(A)  # noqa: E731
        desc = f"2D dense 64x64, m={m}, isotropic TV, maxit {maxit}"
    f64 = workload == "ring-f64"
    bt = ctx.tensor(b, torch.float64 if f64 else torch.float32)
    for _ in range(args.warmup):
        ctx.solve(plan, bt, dims, opts, f64=f64)
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        U, info, _ = ctx.solve(plan, bt, dims, opts, Ut=x, f64=f64)
    torch.cuda.synchronize()
    t_steps = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile(False)
    args.one_step = lambda: ctx.solve(plan, bt, dims, opts, Ut=x, f64=f64)
    ops, roof = roofline(prof, kern, workload, plan, t_steps, args.steps)
    if workload != "hemisphere":
        roof["cache_resident"] = True  # operator and vectors fit the 256 MB last-level cache
    if workload == "dense-layers":
        # GEMMs over the layers on FP64 MFMA: 2 L m n flops per product
        fl = {k: prof[k][3] * 2.0 * L * m * (q * p) for k in prof}
        for k in ops:
            ops[k]["TFLOPs"] = fl[k] / (prof[k][0] * 1e-3) / 1e12 if prof[k][0] > 0 else None
        dom = max(ops, key=lambda k: prof[k][0])
        hbm = dict(roof)
        roof = {"bound": "mfma", "kernel": kern[dom], "achieved": ops[dom]["TFLOPs"],
                "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": ops[dom]["TFLOPs"] / FP64_MFMA_PEAK_TFLOPS,
                "traffic": hbm.get("traffic"), "flops_per_launch": fl[dom] / max(prof[dom][1], 1),
                "ms_per_launch": ops[dom]["ms_per_launch"], "precision": "fp64 MFMA (exact products)",
                "hbm": {"achieved": hbm["achieved"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                        "frac": hbm["frac"], "bytes_per_launch": hbm["bytes_per_launch"]},
                "products_GBps_over_wall": hbm.get("products_GBps_over_wall"),
                "cache_resident": True}
    cpu = None
    if not args.no_cpu:
        op = op_fn()
        solve = O.solve_f64 if f64 else O.solve
        if workload == "dense-layers" and L > 8:
            # the CPU solve is linear in the layer count: time 8 layers, scale
            Ls = 8
            bs8 = b[:Ls * m]
            runs, el, thr = cpu_time(lambda: solve(op, bs8, (q, Ls * p, 1), opts), args.cpu_seconds, 5)
            cpu = {"value": runs / el * Ls / L, "unit": "recons/s", "cores": thr, "kind": "port",
                   "sample": f"{runs} full reconstructions of {Ls} of the {L} layers ({el:.1f} s, "
                             f"C++/OpenMP oracle), scaled by {Ls}/{L}"}
        else:
            runs, el, thr = cpu_time(lambda: solve(op, b, dims, opts), args.cpu_seconds, 5)
            cpu = {"value": runs / el, "unit": "recons/s", "cores": thr, "kind": "port",
                   "sample": f"{runs} full reconstructions ({el:.1f} s, C++/OpenMP oracle)"}
    extra = {"images_per_s": L * args.steps / t_steps} if workload == "dense-layers" else {}
    return {
        **extra,
        "value": args.steps / t_steps, "unit": "recons/s", "steps": args.steps,
        "ms_per_step": 1e3 * t_steps / args.steps, "dtype": "f64" if f64 else "f32",
        "data": "synthetic (seeded phantom, noise)",
        "config": {"workload": desc, "batch": 1, "maxit": maxit, "nnz": plan.nnz},
        "iters": info.total_iters, "rel_error_vs_truth": info.rel_error,
        "operators": ops, "roofline": roof, "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ring",
                    choices=["ring", "dense-batch", "hemisphere", "dense", "dense-layers", "ring-f64",
                             "diff"])
    ap.add_argument("--batch", type=int, default=None)
    ap.add_argument("--layers", type=int, default=64,
                    help="dense-layers: images sharing the one A")
    ap.add_argument("--plain-runs", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the FP64 oracle solve used for the error report")
    ap.add_argument("--grid", type=int, default=None)
    ap.add_argument("--maxit", type=int, default=None)
    ap.add_argument("--f64-sample", type=int, default=3,
                    help="dense-batch: smallest instances also checked against the CPU oracle")
    ap.add_argument("--fp64-sample-ring", type=int, default=16,
                    help="ring: instances compared with their FP64 (GPU) solves")
    ap.add_argument("--fp64-chunk-gib", type=float, default=48.0,
                    help="dense-batch: FP64 copies of A per chunk of the FP64 batch solves")
    ap.add_argument("--vectors", default="f32", choices=["f32", "f64"],
                    help="dense-batch: vector and scalar precision on the bf16 A")
    ap.add_argument("--graph", type=int, default=0, choices=[0, 1],
                    help="1: replay each pair of solver steps from a hipGraph")
    ap.add_argument("--slots", default=None, metavar="FILE",
                    help="after the timed steps, run one more step with every launch "
                         "timed and write the per-slot report to FILE")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run a device-to-device copy of this many MiB "
                         "(calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the "
              "multi-GPU sharding of this path is design-only.", file=sys.stderr)
        sys.exit(2)
    args.batch_set = args.batch is not None
    args.grid_set = args.grid is not None
    args.maxit_set = args.maxit is not None
    if args.batch is None:
        # seven 64-lane groups: 669 recons/s against 656 at 512, 668 at 384
        # and 660 at 320 (two runs each on one box)
        args.batch = 448
    if args.grid is None:
        args.grid = 256
    if args.maxit is None:
        args.maxit = 300 if args.workload == "dense-batch" else 200

    import torch
    import tval3_amd
    ctx = tval3_amd.Context(0)
    if args.workload == "ring":
        out = run_ring(args, ctx)
    elif args.workload == "dense-batch":
        out = run_dense_batch(args, ctx)
    elif args.workload == "diff":
        out = run_diff(args, ctx)
    else:
        out = run_single(args, ctx, args.workload)
    if args.slots:
        # outside the timed region: an event pair per launch slows the step
        ctx.profile(2)
        args.one_step()
        report = ctx.profile_report()
        ctx.profile(False)
        with open(args.slots, "w") as f:
            f.write(f"per-slot launch times of one step of --workload {args.workload} "
                    f"({out['config']['workload']})\n{report}")
        log(report)
    if args.copy_mb:
        src = torch.ones(args.copy_mb * 262144, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
    line = {"metric": METRIC, "value": out.pop("value"), "unit": out.pop("unit"),
            "n_gpus": 1, "steps": out.pop("steps"), "warmup": args.warmup,
            "ms_per_step": out.pop("ms_per_step"), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": out.pop("dtype"),
            "data": out.pop("data"), "config": out.pop("config")}
    line.update(out)
    cpu = line.get("cpu_baseline")
    if cpu:
        line["gpu_over_cpu"] = line["value"] / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
