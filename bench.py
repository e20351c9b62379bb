"""Benchmark: nearest-hit Mrays/s on the 2,097,152-triangle heightfield.

One step = one batched nearest-hit pass of all 8,388,608 rays (4M coherent
camera rays + 4M incoherent rays, t_max 2000) through the HIP engine, rays and
scene resident in HBM.  `value` is the plain pooled walk in the default hit
mode (SQ_HITS_REFERENCE: the reference's hits on every ray, its walk order).
Reported beside it: the guarded near-first walk (SQ_HITS_GUARDED) with the
number of rays where its answer differs from the reference's, the
learned-proposal leg (train the MLP on 16M fresh rays, then fused inference +
seeded walk on the same 8M rays; the seeded walk answers in the guarded mode)
and the CPU baseline (all host threads, and one).  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-sample N] [--no-proposal]

`--f64` measures the reference's `f64` feature instead (one JSON line of its
own): nearest-hit rays over 1,000,000 double-precision balls (D = 3), with the
CPU baseline (the f64 restatement of raycast_nearest_batch_parallel) on a
sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "spatial-query_amd"))

METRIC = "Mrays/s nearest-hit, 2M-tri scene: learned-proposal vs plain kernel, exact hits"
HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, published
BF16_PEAK_TFLOPS = 2500.0   # MI355X dense BF16 matrix peak, published


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=-1, help="rays in the CPU baseline sample (-1 = all, 0 = skip)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--rays-per-lane", type=int, default=-1, help="pooled kernel pool size per lane (0 = one ray per thread)")
    ap.add_argument("--lds-stack", type=int, default=0)
    ap.add_argument("--ray-order", type=int, default=-1, help="1: Morton ray order, 0: as given, -1: default")
    ap.add_argument("--descend-min", type=int, default=-1, help="pause descents below this many lanes (-1: default)")
    ap.add_argument("--wide", type=int, default=-1, help="1: walk the 4-wide records, 0: the pair nodes, -1: default")
    ap.add_argument("--packet", type=int, default=-1, help="1: one wave-coherent walk per 64 rays, 0: one ray per lane, -1: default")
    ap.add_argument("--quick", action="store_true", help="short run for profiling: no CPU baseline, no side figures")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a 1 GiB device copy (the known-size kernel that calibrates HBM counters)")
    ap.add_argument("--no-proposal", action="store_true", help="skip the learned-proposal leg")
    ap.add_argument("--proposal-epochs", type=int, default=4)
    ap.add_argument("--proposal-train-rays", type=int, default=1 << 24)
    ap.add_argument("--f64", action="store_true", help="measure the f64 feature (1M balls, 4M rays) instead")
    ap.add_argument("--hit-mode", choices=["reference", "guarded"], default="reference",
                    help="hit mode of the timed walk (reference: the reference's hits, its walk order; guarded: "
                         "box-guarded hits, near-first)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU figures are not produced: rays are "
                                   "independent, so N GPUs would each take a contiguous ray range of a replicated scene",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.f64:
        f64_run(args)
        return

    import numpy as np
    import torch
    import sq_amd
    from sq_amd import workloads as W

    t0 = time.time()
    sc, coh, inc = W.c1()
    rays_h = np.concatenate([coh, inc])
    n = rays_h.shape[0]
    log(f"generated C1: {sc.prims.shape[0]} tris, {n} rays in {time.time() - t0:.1f}s")

    ctx = sq_amd.Context(0)
    tun = ctx.tuning()
    if args.block:
        tun["block"] = args.block
    if args.rays_per_lane >= 0:
        tun["rays_per_lane"] = args.rays_per_lane
    if args.lds_stack:
        tun["lds_stack"] = args.lds_stack
    if args.ray_order >= 0:
        tun["ray_order"] = args.ray_order
    if args.descend_min >= 0:
        tun["descend_min"] = args.descend_min
    if args.wide >= 0:
        tun["wide"] = args.wide
    if args.packet >= 0:
        tun["packet"] = args.packet
    ctx.set_tuning(**tun)
    ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED if args.hit_mode == "guarded" else sq_amd.SQ_HITS_REFERENCE)
    t0 = time.time()
    tree = sq_amd.Tree.build(3, sq_amd.TRIANGLE, sc.prims)
    t_build = time.time() - t0
    bvh = sq_amd.DeviceBvh(ctx, tree, sq_amd.TRIANGLE, sc.prims)
    st = bvh.stats()
    log(f"host build {t_build:.2f}s, nodes {st['node_count']}, depth {st['max_depth']}")

    rays = torch.from_numpy(rays_h).cuda()
    leaf = torch.empty(n, dtype=torch.int32, device="cuda")
    toi = torch.empty(n, dtype=torch.float32, device="cuda")

    def step():
        bvh.raycast_nearest_into(rays, 2000.0, leaf, toi)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        kms.append(ctx.last_kernel_ms())     # HIP events on the engine stream around the launch
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = n / (ms_per_step * 1e-3) / 1e6
    k_ms = float(np.mean(kms))
    log(f"nearest: {ms_per_step:.3f} ms/step, kernel {k_ms:.3f} ms, {value:.1f} Mrays/s")

    # Algorithmic bytes: B_ray = 8D + 8 + (8D+8) N_node + (P+4) N_prim, D = 3, P = 36,
    # N_node / N_prim the order-independent minimal work at the final toi.
    tl = torch.where(leaf >= 0, toi, torch.full_like(toi, 2000.0))
    nodes, prims = bvh.required_work(rays, tl)
    ctx.synchronize()
    n_node = nodes.double().sum().item()
    n_prim = prims.double().sum().item()
    bytes_launch = n * (24 + 8) + 32 * n_node + 40 * n_prim
    achieved = bytes_launch / (k_ms * 1e-3) / 1e9
    # HBM traffic per launch and the VALU figures of the timed kernel, from its
    # rocprofv3 counter runs (profiles/k_nearest_pool_c1_heightfield.md): fetch
    # and write bytes scaled by the copy-kernel calibration.
    traffic, valu = None, None
    tpath = os.path.join(ROOT, "profiles", "traffic_c1_nearest.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            prof = json.load(fh)
        tun_now = ctx.tuning()
        if (prof.get("hit_mode", "reference") == args.hit_mode and prof.get("wide", 0) == tun_now["wide"]
                and prof.get("lookahead", 0) == tun_now["lookahead"]
                and prof.get("lookahead_pause", 24) == tun_now["lookahead_pause"]):
            traffic = int(prof["traffic_bytes"])
            valu = prof.get("valu")
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 heightfield and rays)",
        "method": ("plain pooled walk, the reference's order (exact: the reference's hit on every ray)"
                   if args.hit_mode == "reference" else "plain pooled near-first walk over guarded hits"),
        "config": {"workload": "C1 heightfield: 1025x1025 grid, 2,097,152 triangles; 4,194,304 coherent + "
                               "4,194,304 incoherent rays; nearest-hit, t_max 2000",
                   "rays_per_step": n, "tree_nodes": st["node_count"], "tree_depth": st["max_depth"],
                   "hit_mode": args.hit_mode, "tuning": ctx.tuning()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel_ms": round(k_ms, 4), "algorithmic_bytes_per_launch": int(bytes_launch),
                     "hbm_measured_gbs": round(traffic / (k_ms * 1e-3) / 1e9, 1) if traffic else None,
                     "hbm_measured_frac": round(traffic / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                     "valu": valu,
                     "note": "achieved: the algorithmic bytes (every node box and triangle of the order-independent "
                             "minimal walk read once, rays in, results out) over the kernel time; "
                             "hbm_measured: FETCH_SIZE + WRITE_SIZE bytes of the same kernel over the same time "
                             "(the scene is mostly served from L2 and the Infinity Cache); valu: the share of SIMD "
                             "cycles issuing VALU and the active-lane share of those issues (the walk's limit)",
                     "mean_nodes_per_ray": round(n_node / n, 2), "mean_prims_per_ray": round(n_prim / n, 2)},
    }

    if not args.quick:
        # side figures: coherent / incoherent halves, any-hit
        half = n // 2
        for name, sl in (("coherent", slice(0, half)), ("incoherent", slice(half, n))):
            rr = rays[sl].contiguous(); l2 = leaf[sl].contiguous(); t2 = toi[sl].contiguous()
            bvh.raycast_nearest_into(rr, 2000.0, l2, t2)
            ts = []
            for _ in range(3):
                bvh.raycast_nearest_into(rr, 2000.0, l2, t2)
                ts.append(ctx.last_kernel_ms())
            out[f"nearest_{name}_mrays"] = round(half / (min(ts) * 1e-3) / 1e6, 1)
        # the coherent half as a coherent batch: tuning `packet` (the batch in its
        # own Morton order, one wave-coherent walk per 64 rays); same answers
        tun0 = ctx.tuning()
        rr = rays[:half].contiguous()
        pl = torch.empty(half, dtype=torch.int32, device="cuda"); pt = torch.empty(half, device="cuda")
        ctx.set_tuning(**dict(tun0, packet=1))
        bvh.raycast_nearest_into(rr, 2000.0, pl, pt)
        ts = []
        for _ in range(5):
            bvh.raycast_nearest_into(rr, 2000.0, pl, pt)
            ts.append(ctx.last_kernel_ms())
        ctx.set_tuning(**tun0)
        same = bool(torch.equal(pl, leaf[:half])) and bool(torch.equal(pt.view(torch.int32), toi[:half].view(torch.int32)))
        out["nearest_coherent_packet"] = {
            "mrays": round(half / (min(ts) * 1e-3) / 1e6, 1), "ms": round(min(ts), 4),
            "speedup_vs_one_ray_per_lane": round(out["nearest_coherent_mrays"] and
                                                 half / (min(ts) * 1e-3) / 1e6 / out["nearest_coherent_mrays"], 3),
            "same_answers": same,
            "note": "sort (batch-bounds Morton key, rocPRIM radix sort) + packet walk, both inside the timed events"}
        log(f"coherent half, packet walk: {out['nearest_coherent_packet']['mrays']} Mrays/s "
            f"({out['nearest_coherent_packet']['speedup_vs_one_ray_per_lane']}x one ray per lane), same answers {same}")
        anyh = torch.empty(n, dtype=torch.uint8, device="cuda")
        bvh.raycast_any(rays, 2000.0, out=anyh)
        ts = []
        for _ in range(3):
            bvh.raycast_any(rays, 2000.0, out=anyh)
            ts.append(ctx.last_kernel_ms())
        out["anyhit_mrays"] = round(n / (min(ts) * 1e-3) / 1e6, 1)
        out["hit_fraction"] = round(float((leaf >= 0).float().mean().item()), 4)
        # the same step end to end from host memory: rays copied in from pinned
        # host memory, the walk, leaf and toi copied back to pinned host memory
        # (H2D + kernel + D2H on the engine's stream; reported beside value,
        # which starts with the rays resident in HBM)
        rays_pin = torch.from_numpy(rays_h).pin_memory()
        leaf_pin = torch.empty(n, dtype=torch.int32).pin_memory()
        toi_pin = torch.empty(n, dtype=torch.float32).pin_memory()
        rays_d = torch.empty_like(rays)
        l_h = torch.empty_like(leaf); t_h = torch.empty_like(toi)

        def host_step():
            rays_d.copy_(rays_pin, non_blocking=True)
            bvh.raycast_nearest_into(rays_d, 2000.0, l_h, t_h)
            leaf_pin.copy_(l_h, non_blocking=True)
            toi_pin.copy_(t_h, non_blocking=True)

        host_step()
        ctx.synchronize()
        torch.cuda.synchronize()
        ts = []
        for _ in range(10):
            t0 = time.perf_counter()
            host_step()
            ctx.synchronize()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        same_h = bool(torch.equal(leaf_pin, leaf.cpu())) and bool(torch.equal(toi_pin.view(torch.int32), toi.cpu().view(torch.int32)))
        med = float(np.median(ts))
        out["with_host_transfers"] = {"mrays": round(n / (med * 1e-3) / 1e6, 1), "ms": round(med, 4),
                                      "h2d_bytes": int(n * 24), "d2h_bytes": int(n * 8), "same_answers": same_h,
                                      "note": "median of 10: rays from pinned host memory, the walk, leaf and toi back "
                                              "to pinned host memory; value starts with the rays resident in HBM"}
        log(f"with host transfers: {out['with_host_transfers']['mrays']} Mrays/s ({med:.3f} ms), same answers {same_h}")
        # the other hit mode on the same rays: its speed, and the rays whose
        # answer (id or toi bits) differs from this run's
        other = "guarded" if args.hit_mode == "reference" else "reference"
        ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED if other == "guarded" else sq_amd.SQ_HITS_REFERENCE)
        lo = torch.empty_like(leaf); to = torch.empty_like(toi)
        bvh.raycast_nearest_into(rays, 2000.0, lo, to)
        ts = []
        for _ in range(5):
            bvh.raycast_nearest_into(rays, 2000.0, lo, to)
            ts.append(ctx.last_kernel_ms())
        ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED if args.hit_mode == "guarded" else sq_amd.SQ_HITS_REFERENCE)
        differ = int(((lo != leaf) | (to.view(torch.int32) != toi.view(torch.int32))).sum().item())
        out[f"{other}_mode"] = {"mrays": round(n / (float(np.median(ts)) * 1e-3) / 1e6, 1),
                                "kernel_ms": round(float(np.median(ts)), 4),
                                "rays_whose_answer_differs": differ, "of_rays": n}
        log(f"{other} mode: {out[other + '_mode']['mrays']} Mrays/s, {differ} of {n} rays answer differently")

    if not args.no_proposal and not args.quick:
        out["learned_proposal"] = proposal_leg(args, ctx, bvh, sc, rays, leaf, toi, k_ms, n)
        out["learned_proposal"]["instanced"] = proposal_instanced_leg(args, ctx)
        lp = out["learned_proposal"]
        out["method"] += (f"; the learned-proposal path (fused BF16 MFMA inference + seeded walk, guarded hits) "
                          f"measures {lp['mrays']} Mrays/s = {lp['speedup_vs_value']}x this walk and "
                          f"{lp['speedup_vs_guarded_plain']}x the guarded plain walk, and a perfect proposal (t_max = "
                          f"the exact toi per ray) bounds any proposal at {lp['perfect_limit_bound']}x the guarded walk")

    if args.cpu_sample and not args.quick:
        import oracle
        m = n if args.cpu_sample < 0 else min(args.cpu_sample, n)
        idx = np.linspace(0, n - 1, m).astype(np.int64)       # evenly spread over both halves
        sample = np.ascontiguousarray(rays_h[idx])
        cores = os.cpu_count() or 1                            # every host thread (nproc)
        o = oracle.Scene(3, oracle.TRI, sc.prims, guarded=args.hit_mode == "guarded")
        t0 = time.perf_counter()
        ref = o.raycast_nearest(sample, 2000.0, threads=cores)
        cpu_s = time.perf_counter() - t0
        got_leaf = leaf.cpu().numpy()[idx].astype(np.int64)
        got_t = toi.cpu().numpy()[idx]
        agree = (got_leaf == ref["leaf"]) & (got_t.view(np.uint32) == ref["t"].view(np.uint32))
        # one thread on a bounded sample (every 32nd ray of the same sample)
        one = np.ascontiguousarray(sample[::32])
        t0 = time.perf_counter()
        o.raycast_nearest(one, 2000.0, threads=1)
        one_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(m / cpu_s / 1e6, 4), "unit": "Mrays/s", "cores": cores, "kind": "port",
                               "sample": f"{m} of the {n} rays, evenly spaced over both halves; C++ restatement of "
                                         f"raycast_nearest_batch_parallel (contiguous chunks, one per thread, "
                                         f"{cores} threads = nproc), built -O3 -ffp-contract=off",
                               "seconds": round(cpu_s, 2), "cpu_model": _cpu_model(),
                               "one_thread": {"value": round(len(one) / one_s / 1e6, 4), "unit": "Mrays/s",
                                              "rays": int(len(one)), "seconds": round(one_s, 2)}}
        out["id_agreement_on_sample"] = float(agree.mean())
        log(f"cpu baseline {m / cpu_s / 1e6:.3f} Mrays/s on {cores} threads, {len(one) / one_s / 1e6:.3f} on one; "
            f"agreement {agree.mean():.6f}")
    if args.copy_calib:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        out["copy_calib_bytes"] = {"read": a.numel() * 4, "written": b.numel() * 4}
    print(json.dumps(out), flush=True)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def f64_run(args):
    """The f64 feature: nearest hits of 4,194,304 rays (origins U[-1000,1000]^3,
    unit directions, t_max 4000) over 1,000,000 balls (centres U[-1000,1000]^3,
    radii U[0.5,4]) in double.  Kernel time from HIP events on the query stream."""
    import numpy as np
    import torch
    import sq_amd

    rng = np.random.default_rng(0xF64B)
    nb, n, max_toi = 1_000_000, 1 << 22, 4000.0
    balls = np.concatenate([rng.uniform(-1000, 1000, (nb, 3)), rng.uniform(0.5, 4.0, (nb, 1))], 1)
    o = rng.uniform(-1000, 1000, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays_h = np.concatenate([o, d], 1)
    ctx = sq_amd.Context(0)
    b = sq_amd.DeviceBvh64(ctx, 3, sq_amd.SPHERE, balls)
    rays = torch.from_numpy(rays_h).cuda()
    res = {}
    for _ in range(args.warmup):
        res["w"] = b.raycast_nearest(rays, max_toi)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        res["w"] = b.raycast_nearest(rays, max_toi)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    leaf = res["w"]["leaf"].cpu().numpy().astype(np.int64)
    toi = res["w"]["toi"].cpu().numpy()
    out = {"metric": "Mrays/s nearest-hit, f64 feature: 1M balls (D = 3), exact hits", "value": round(n / ms / 1e3, 2),
           "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded numpy PCG64 balls and rays)",
           "config": {"workload": "f64: 1,000,000 balls U[-1000,1000]^3 r U[0.5,4]; 4,194,304 rays, t_max 4000",
                      "rays_per_step": n, "tree": b.stats()}}
    if args.cpu_sample and not args.quick:
        import oracle
        m = n if args.cpu_sample < 0 else min(args.cpu_sample, n)
        idx = np.linspace(0, n - 1, m).astype(np.int64)
        cores = os.cpu_count() or 1
        ref = oracle.Scene64(3, oracle.SPHERE, balls)
        t0 = time.perf_counter()
        w = ref.bvh_raycast_nearest(rays_h[idx], max_toi, threads=cores)
        cpu_s = time.perf_counter() - t0
        agree = (leaf[idx] == w["leaf"]) & (toi[idx].view(np.uint64) == w["t"].view(np.uint64))
        out["cpu_baseline"] = {"value": round(m / cpu_s / 1e6, 4), "unit": "Mrays/s", "cores": cores, "kind": "port",
                               "sample": f"{m} of the {n} rays, evenly spaced; f64 restatement of "
                                         f"raycast_nearest_batch_parallel (contiguous chunks, {cores} threads)",
                               "seconds": round(cpu_s, 2)}
        out["id_agreement_on_sample"] = float(agree.mean())
        # algorithmic bytes per ray from the reference-order walk's counts on the sample: ray in (48 B), leaf + toi
        # out (12 B), one 112-byte pair record per two node boxes tested, 36 B per ball tested (32 B + id)
        nodes, prims = w["nodes_tested"] / m, w["prims_tested"] / m
        per_ray = 48 + 12 + 112 * nodes / 2 + 36 * prims
        achieved = per_ray * n / (ms * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                           "note": "bytes per ray estimated from the reference-order walk's node and primitive tests "
                                   "on the sample; the device's near-first walk tests a different set",
                           "nodes_tested_per_ray": round(nodes, 2), "prims_tested_per_ray": round(prims, 2)}
        log(f"f64: {out['value']} Mrays/s; cpu {m / cpu_s / 1e6:.3f} Mrays/s; agreement {agree.mean():.6f}")
    b.close()
    print(json.dumps(out), flush=True)


def proposal_leg(args, ctx, bvh, sc, rays, leaf, toi, plain_ms, n):
    """Learned subtree proposal on the same rays: train on fresh draws from the
    C1 distributions (jittered camera rays of the same camera, incoherent rays
    of another seed), then time fused inference + the seeded exact walk."""
    import numpy as np
    import torch
    import sq_amd
    from sq_amd import workloads as W

    res = {}
    mode0 = ctx.hit_mode()
    # the seeded walk answers under SQ_HITS_GUARDED (order-independent hits), so
    # its exactness and speed are measured against the guarded plain walk
    ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED)
    gl = torch.empty_like(leaf); gt = torch.empty_like(toi)
    bvh.raycast_nearest_into(rays, 2000.0, gl, gt)
    ts = []
    for _ in range(5):
        bvh.raycast_nearest_into(rays, 2000.0, gl, gt)
        ts.append(ctx.last_kernel_ms())
    guarded_ms = float(np.median(ts))
    # bound: the guarded plain walk with t_max = the exact toi of each ray (what
    # a perfect, free proposal could give)
    tl = torch.where(gl >= 0, gt, torch.full_like(gt, 2000.0)).contiguous()
    l2 = torch.empty_like(leaf); t2 = torch.empty_like(toi)
    bvh.raycast_nearest_into(rays, 2000.0, l2, t2, max_toi_per_ray=tl)
    ts = []
    for _ in range(5):
        bvh.raycast_nearest_into(rays, 2000.0, l2, t2, max_toi_per_ray=tl)
        ts.append(ctx.last_kernel_ms())
    res["guarded_plain_ms"] = round(guarded_ms, 4)
    res["perfect_limit_ms"] = round(min(ts), 4)
    res["perfect_limit_bound"] = round(guarded_ms / min(ts), 3)

    p = sq_amd.Proposal(bvh, cut_depth=12, rff_scale=1.0, seed=0x5EED)
    t0 = time.time()
    train = W.proposal_training_rays(sc, args.proposal_train_rays)
    rt = torch.from_numpy(train).cuda()
    lab = p.labels(bvh.raycast_nearest(rt, 2000.0)["leaf"])
    torch.cuda.synchronize()
    t_data = time.time() - t0
    # training throughput: one timed epoch-sized pass after the first
    torch.cuda.synchronize(); t0 = time.time()
    losses = p.fit(rt, lab, epochs=args.proposal_epochs, batch=65536, lr=1e-3, seed=1)
    torch.cuda.synchronize()
    t_train = time.time() - t0
    steps = args.proposal_epochs * (rt.shape[0] // 65536)
    flop_ray = 2.0 * (6 * 64 + 128 * 256 + 2 * 256 * 256 + 256 * p.n_subtrees)
    res.update(subtrees=p.n_subtrees, train_rays=int(rt.shape[0]), epochs=args.proposal_epochs,
               data_s=round(t_data, 2), train_s=round(t_train, 2), loss_first=round(losses[0], 4),
               loss_last=round(losses[-1], 4),
               train_tflops=round(3 * flop_ray * 65536 * steps / t_train / 1e12, 1))
    del rt, lab, train

    top = p.infer(rays)
    pl = torch.empty_like(leaf); pt = torch.empty_like(toi)
    p.raycast_nearest_into(rays, 2000.0, top, pl, pt)
    torch.cuda.synchronize()
    exact = bool(torch.equal(pl, gl)) and bool(torch.equal(pt.view(torch.int32), gt.view(torch.int32)))
    vs_value = int(((pl != leaf) | (pt.view(torch.int32) != toi.view(torch.int32))).sum().item())
    ti, tw = [], []
    for _ in range(5):
        p.infer(rays, out=top); ti.append(ctx.last_kernel_ms())
        p.raycast_nearest_into(rays, 2000.0, top, pl, pt); tw.append(ctx.last_kernel_ms())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        p.infer(rays, out=top)
        p.raycast_nearest_into(rays, 2000.0, top, pl, pt)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    lab_e = p.labels(gl).cpu().numpy()
    t = top.cpu().numpy().view(np.uint16).astype(np.int64)
    half = n // 2
    def rate(sl):
        ok = lab_e[sl] >= 0
        return round(float((t[sl][ok] == lab_e[sl][ok, None]).any(axis=1).mean()), 4)
    cnt = p.raycast_nearest(rays, 2000.0, top, count=True)
    plain_cnt = p.raycast_nearest(rays, 2000.0, None, count=True)
    infer_ms = float(np.median(ti))
    res.update(mrays=round(n / (ms * 1e-3) / 1e6, 2), ms_per_step=round(ms, 4),
               speedup_vs_value=round(plain_ms / ms, 3), speedup_vs_guarded_plain=round(guarded_ms / ms, 3),
               exact_vs_guarded_plain=exact, rays_differing_from_value=vs_value,
               infer_ms=round(infer_ms, 4), seeded_walk_ms=round(float(np.median(tw)), 4),
               infer_tflops=round(flop_ray * n / (infer_ms * 1e-3) / 1e12, 1),
               mfma_utilisation=round(flop_ray * n / (infer_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4),
               top4_hit_rate_coherent=rate(slice(0, half)), top4_hit_rate_incoherent=rate(slice(half, n)),
               seeded_fraction=round(cnt["seeded"] / n, 4),
               nodes_per_ray_plain=round(plain_cnt["nodes"] / n, 2), nodes_per_ray_seeded=round(cnt["nodes"] / n, 2),
               prims_per_ray_plain=round(plain_cnt["prims"] / n, 2), prims_per_ray_seeded=round(cnt["prims"] / n, 2),
               flop_per_ray=flop_ray)
    log(f"proposal: {res['mrays']} Mrays/s ({res['speedup_vs_guarded_plain']}x guarded plain, {res['speedup_vs_value']}x "
        f"value), infer {res['infer_ms']} ms at {res['infer_tflops']} TFLOP/s, walk {res['seeded_walk_ms']} ms, exact vs "
        f"guarded {exact}, bound {res['perfect_limit_bound']}x")
    p.close()
    ctx.set_hit_mode(mode0)
    return res


def proposal_instanced_leg(args, ctx):
    """The proposal's instanced half: the two-level scene of 16 noise-displaced
    icospheres (2k-50k triangles) placed as 4,096 instances (uniform rotations,
    translations U[-1000,1000]^3, scales U[0.5,4]); the top tree over the
    instances cut at depth 12; trained on fresh random rays (origins
    U[-1500,1500]^3, uniform directions, t_max 4000) labelled by the exact
    two-level walk, evaluated on 8M rays of another seed.  All walks answer
    under SQ_HITS_GUARDED (the seeded walk's mode)."""
    import numpy as np
    import torch
    import sq_amd
    from sq_amd import workloads as W

    mode0 = ctx.hit_mode()
    ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED)
    tmax = 4000.0
    sc = W.instanced_scene()
    meshes = [sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, m) for m in sc["meshes"]]
    tl = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.MESH, sc["rot"], sc["trans"], meshes=meshes, mesh_of=sc["mesh_of"],
                     scale=sc["scale"])
    n = 1 << 23
    rays = torch.from_numpy(W.random_rays(n, 0x1C3E_0000_0001, -1500.0, 1500.0)).cuda()
    res = dict(scene="4,096 instances of 16 meshes (2k-50k triangles), rays U[-1500,1500]^3 x uniform directions, "
                     "t_max 4000", eval_rays=n)

    def timed(fn, reps=5):
        fn()
        ts = []
        for _ in range(reps):
            fn()
            ts.append(ctx.last_kernel_ms())
        return float(np.median(ts))

    plain = tl.raycast_nearest(rays, tmax, sub=True)
    plain_ms = timed(lambda: tl.raycast_nearest(rays, tmax))
    hit = plain["leaf"] >= 0
    lim = torch.where(hit, plain["toi"], torch.full_like(plain["toi"], tmax)).contiguous()
    perfect_ms = timed(lambda: tl.raycast_nearest(rays, tmax, max_toi_per_ray=lim))
    res.update(hit_fraction=round(float(hit.float().mean().item()), 4), guarded_plain_ms=round(plain_ms, 4),
               guarded_plain_mrays=round(n / (plain_ms * 1e-3) / 1e6, 1), perfect_limit_ms=round(perfect_ms, 4),
               perfect_limit_bound=round(plain_ms / perfect_ms, 3))

    p = sq_amd.Proposal(tl, cut_depth=12, rff_scale=1.0, seed=0x5EED)
    t0 = time.time()
    rt = torch.from_numpy(W.random_rays(args.proposal_train_rays, 0x1C3E_0000_0002, -1500.0, 1500.0)).cuda()
    lab = p.labels(tl.raycast_nearest(rt, tmax)["leaf"])
    torch.cuda.synchronize()
    t_data = time.time() - t0
    t0 = time.time()
    losses = p.fit(rt, lab, epochs=args.proposal_epochs, batch=65536, lr=1e-3, seed=1)
    torch.cuda.synchronize()
    res.update(subtrees=p.n_subtrees, train_rays=int(rt.shape[0]), labelled_train_rays=int((lab >= 0).sum().item()),
               epochs=args.proposal_epochs, data_s=round(t_data, 2), train_s=round(time.time() - t0, 2),
               loss_first=round(losses[0], 4), loss_last=round(losses[-1], 4))
    del rt, lab

    top = p.infer(rays)
    got = p.raycast_nearest(rays, tmax, top, count=True)
    torch.cuda.synchronize()
    exact = all(bool(torch.equal(got[k], plain[k])) for k in ("leaf", "sub")) and \
        bool(torch.equal(got["toi"].view(torch.int32), plain["toi"].view(torch.int32)))
    infer_ms = timed(lambda: p.infer(rays, out=top))
    pl = torch.empty_like(plain["leaf"]); pt = torch.empty_like(plain["toi"])
    walk_ms = timed(lambda: p.raycast_nearest_into(rays, tmax, top, pl, pt))
    lab_e = p.labels(plain["leaf"]).cpu().numpy()
    tt = top.cpu().numpy().view(np.uint16).astype(np.int64)
    ok = lab_e >= 0
    learned_ms = infer_ms + walk_ms
    res.update(exact_vs_guarded_plain=exact, infer_ms=round(infer_ms, 4), seeded_walk_ms=round(walk_ms, 4),
               mrays=round(n / (learned_ms * 1e-3) / 1e6, 1), speedup_vs_guarded_plain=round(plain_ms / learned_ms, 3),
               seeded_walk_speedup=round(plain_ms / walk_ms, 3),
               top4_hit_rate=round(float((tt[ok] == lab_e[ok, None]).any(axis=1).mean()), 4),
               seeded_fraction=round(got["seeded"] / n, 4))
    log(f"proposal (instanced): hits {res['hit_fraction']}, bound {res['perfect_limit_bound']}x, plain {plain_ms:.3f} ms, "
        f"infer {infer_ms:.3f} + walk {walk_ms:.3f} ms, top-4 hit rate {res['top4_hit_rate']}, exact {exact}")
    p.close()
    tl.close()
    for m in meshes:
        m.close()
    ctx.set_hit_mode(mode0)
    return res


if __name__ == "__main__":
    main()
