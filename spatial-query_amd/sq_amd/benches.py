"""Workload drivers for the batched/instanced scene and the N-dimensional scene.

    python -m sq_amd.benches c2 [--frames 60] [--batches 256] [--no-tune]
    python -m sq_amd.benches c4
    python -m sq_amd.benches build
    python -m sq_amd.benches consumer
    python -m sq_amd.benches dispatch
    python -m sq_amd.benches f64

c2: 16 noise-displaced icospheres (2k-50k triangles, log-uniform) placed as
4,096 instances (uniform rotations, translations U[-1000,1000]^3, scales
U[0.5,4]) plus 100,000 boxes moving with constant velocities U[-5,5]^3.  Each
frame refits the boxes on the device and answers 256 nearest-hit batches
(sizes log-uniform in 1k-1M rays, origins U[-1500,1500]^3, uniform directions,
t_max 4000; even batches on the instances, odd ones on the boxes) with two
persistent batch launches, the whole frame recorded once into a graph and
replayed.  The tuner picks block size, pool size and stack depth first.  A
second series rebuilds the boxes on the device every frame instead of
refitting (sq_bvh_rebuild) and replays a graph of the two submits.

c4: 1,000,000 boxes in D = 6 from a 32-component Gaussian mixture; 1M
nearest-hit rays, 262,144 overlap queries (sorted id lists) and 262,144 box
casts (half-extent 0.5, sweep U(0,50]).

Each prints one JSON line with timings and the consistency checks it ran
(every batch against its single-scene call, in full).  The CPU-oracle checks of
these workloads live in tests/test_gpu_workloads.py.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np


def log(msg):
    print(f"[benches] {msg}", file=sys.stderr, flush=True)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


def _same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def c2(frames=60, n_batches=256, tune=True, check=True, seed=0xC2C2_0000_000A):
    import torch
    import sq_amd
    from sq_amd import workloads as W
    from sq_amd.autotune import autotune

    out = dict(workload="c2: 16 meshes x 4096 instances + 100k moving boxes; "
                        f"{frames} frames x {n_batches} batches (log-U 1k-1M rays)")
    t0 = time.time()
    sc = W.instanced_scene()
    s = W.Stream(seed)
    boxes = W.boxes_scene(n=100_000, seed=int(s.u64(1)[0]), lo=-1000.0, hi=1000.0, hmin=0.5, hmax=5.0).prims
    vel = s.range(-5.0, 5.0, 100_000 * 3).reshape(100_000, 3).astype(np.float32)
    sizes = np.round(np.exp(s.range(math.log(1e3), math.log(1e6), n_batches).astype(np.float64))).astype(np.int64)
    ray_seeds = [int(v) for v in s.u64(n_batches)]
    rays_h = [W.random_rays(int(n), rs, -1500.0, 1500.0) for n, rs in zip(sizes, ray_seeds)]
    tmax = 4000.0
    log(f"generated in {time.time() - t0:.1f}s: {int(sizes.sum())} rays per frame")
    out.update(rays_per_frame=int(sizes.sum()), batches=n_batches, frames=frames,
               mesh_tris=[int(m.shape[0]) for m in sc["meshes"]], instances=int(sc["rot"].shape[0]), boxes=100_000)

    ctx = sq_amd.Context(0)
    stream = torch.cuda.Stream()
    ctx.set_stream(stream.cuda_stream)
    t0 = time.time()
    meshes = [sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, m) for m in sc["meshes"]]
    tlas = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.MESH, sc["rot"], sc["trans"], meshes=meshes, mesh_of=sc["mesh_of"],
                       scale=sc["scale"])
    bvh = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.AABB, boxes)
    prims_dev = _dev(boxes)
    rays = [_dev(r) for r in rays_h]
    torch.cuda.synchronize()
    log(f"scenes built in {time.time() - t0:.1f}s")
    qt = [(tlas, rays[i], tmax) for i in range(0, n_batches, 2)]
    qb = [(bvh, rays[i], tmax) for i in range(1, n_batches, 2)]

    def frame_sets():
        return sq_amd.BatchSet(ctx, qt), sq_amd.BatchSet(ctx, qb)

    if tune:
        def prepare():
            bt, bb = frame_sets()

            def go():
                bvh.refit(prims_dev)
                bt.submit()
                bb.submit()
                return tuple(bt.leaf) + tuple(bb.leaf)
            prepare.keep = (bt, bb)
            return go
        t0 = time.time()
        res = autotune(ctx, prepare, space=dict(block=(64, 128, 256), rays_per_lane=(2, 4, 8, 16), lds_stack=(4, 8, 16)),
                       reps=2)
        prepare.keep = None
        out["tuning"] = res["best"]
        out["tuning_ms"] = round(res["ms"], 3)
        out["tuning_default_ms"] = round(res["baseline"], 3)
        log(f"tuned in {time.time() - t0:.1f}s: {res['best']} {res['ms']:.2f} ms (default {res['baseline']:.2f} ms)")
    bt, bb = frame_sets()
    out["chunks"] = bt.info()["chunks"] + bb.info()["chunks"]
    # warm-up (sizes buffers), then record the frame once
    bvh.refit(prims_dev)
    bt.submit()
    bb.submit()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        bvh.refit(prims_dev)
        bt.submit()
        bb.submit()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def replay_ms():
        with torch.cuda.stream(stream):
            e0.record()
            g.replay()
            e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1)

    # the frame-0 state (freshly built trees): graph replay against the same
    # frame issued as one call per batch
    prims_dev.copy_(torch.from_numpy(boxes))
    torch.cuda.synchronize()
    f0 = [replay_ms() for _ in range(5)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        bvh.refit(prims_dev)
        for (scene, r, tm) in qt + qb:
            scene.raycast_nearest(r, tm)
    ctx.synchronize()
    single = (time.perf_counter() - t0) * 1e3 / 3
    f0m = float(np.median(f0))
    out.update(frame0_ms=round(f0m, 3), frame0_mrays_per_s=round(float(sizes.sum()) / (f0m * 1e-3) / 1e6, 1),
               single_calls_frame0_ms=round(single, 3), batching_gain=round(single / f0m, 3))
    log(f"frame 0: graph {f0m:.2f} ms, single calls {single:.2f} ms")
    fms, parity = [], []
    for f in range(frames):
        moved = boxes.copy()
        d = vel * np.float32(f)
        moved[:, :3] += d
        moved[:, 3:] += d
        prims_dev.copy_(torch.from_numpy(moved))
        torch.cuda.synchronize()
        fms.append(replay_ms())
        if check and f in (0, frames - 1):
            # every batch equals its single-scene call
            ok = True
            for i, (scene, r, tm) in enumerate(qt):
                w = scene.raycast_nearest(r, tm)
                ctx.synchronize()
                ok &= _same(bt.leaf[i].cpu(), w["leaf"].cpu()) and _same_bits(bt.toi[i].cpu(), w["toi"].cpu())
            for i, (scene, r, tm) in enumerate(qb):
                w = scene.raycast_nearest(r, tm)
                ctx.synchronize()
                ok &= _same(bb.leaf[i].cpu(), w["leaf"].cpu()) and _same_bits(bb.toi[i].cpu(), w["toi"].cpu())
            parity.append(dict(frame=f, batches_equal_single_calls=bool(ok)))
            log(f"frame {f}: batches == single calls {ok}")
    out["parity"] = parity
    fm = float(np.median(fms))
    out.update(frame_ms_median=round(fm, 3), frame_ms_min=round(float(np.min(fms)), 3),
               mrays_per_s=round(float(sizes.sum()) / (fm * 1e-3) / 1e6, 1),
               frame_ms=[round(v, 2) for v in fms])
    # the refit alone
    with torch.cuda.stream(stream):
        e0.record()
        for _ in range(10):
            bvh.refit(prims_dev)
        e1.record()
    e1.synchronize()
    out["refit_ms"] = round(e0.elapsed_time(e1) / 10, 4)
    log(f"60 frames: median {fm:.2f} ms ({out['mrays_per_s']} Mrays/s; refit-only trees loosen as the boxes "
        f"drift), refit {out['refit_ms']} ms")

    # the same frames with the boxes rebuilt on the device every frame (the
    # tree keeps its arrays, so the batch set and the graph of the two
    # submits stay valid): frame = device rebuild + graph replay
    prims_dev.copy_(torch.from_numpy(boxes))
    dbvh = sq_amd.DeviceBvh.build_device(ctx, 3, sq_amd.AABB, prims_dev)
    bb2 = sq_amd.BatchSet(ctx, [(dbvh, r, tm) for (_, r, tm) in qb])
    bt.submit()
    bb2.submit()
    torch.cuda.synchronize()
    g2 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g2, stream=stream):
        bt.submit()
        bb2.submit()
    rms, rb_ms, parity2 = [], [], []
    for f in range(frames):
        moved = boxes.copy()
        d = vel * np.float32(f)
        moved[:, :3] += d
        moved[:, 3:] += d
        prims_dev.copy_(torch.from_numpy(moved))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dbvh.rebuild(prims_dev)
        t1 = time.perf_counter()
        with torch.cuda.stream(stream):
            g2.replay()
        stream.synchronize()
        t2 = time.perf_counter()
        rb_ms.append((t1 - t0) * 1e3)
        rms.append((t2 - t0) * 1e3)
        if check and f in (0, frames - 1):
            ok = True
            for i, (_, r, tm) in enumerate(qb):
                w = dbvh.raycast_nearest(r, tm)
                ctx.synchronize()
                ok &= _same(bb2.leaf[i].cpu(), w["leaf"].cpu()) and _same_bits(bb2.toi[i].cpu(), w["toi"].cpu())
            parity2.append(dict(frame=f, batches_equal_single_calls=bool(ok)))
            log(f"rebuild frame {f}: batches == single calls {ok}")
    rm = float(np.median(rms))
    out["rebuild_frames"] = dict(frame_ms_median=round(rm, 3), frame_ms_first=round(rms[0], 3),
                                 frame_ms_last=round(rms[-1], 3), rebuild_ms_median=round(float(np.median(rb_ms)), 3),
                                 mrays_per_s=round(float(sizes.sum()) / (rm * 1e-3) / 1e6, 1),
                                 frame_ms=[round(v, 2) for v in rms], parity=parity2)
    log(f"{frames} rebuild frames: median {rm:.2f} ms (rebuild {np.median(rb_ms):.2f} ms), first {rms[0]:.2f}, "
        f"last {rms[-1]:.2f}")

    # the same frames under the refit-or-rebuild policy (sq_bvh_refit_auto:
    # refit, and rebuild on the device once the tree's surface sum passes
    # factor x that of the tree as last built, Tlas::needs_rebuild's test):
    # frame = policy step + graph replay of the two submits; factors 1.1 to 2
    # swept, the fastest reported as policy_frames
    def record():
        b = sq_amd.BatchSet(ctx, [(dbvh, r, tm) for (_, r, tm) in qb])
        bt.submit()
        b.submit()
        torch.cuda.synchronize()
        gg = torch.cuda.CUDAGraph()
        with torch.cuda.graph(gg, stream=stream):
            bt.submit()
            b.submit()
        return b, gg

    bb2.close()
    state = {}

    def policy_run(factor):
        prims_dev.copy_(torch.from_numpy(boxes))
        dbvh.rebuild(prims_dev)
        if state.get("set") is None or not state["set"].valid():
            if state.get("set") is not None:
                state["set"].close()
            state["set"], state["graph"] = record()
        pms, rebuilt, recorded, par = [], [], 0, []
        for f in range(frames):
            moved = boxes.copy()
            d = vel * np.float32(f)
            moved[:, :3] += d
            moved[:, 3:] += d
            prims_dev.copy_(torch.from_numpy(moved))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = dbvh.refit_auto(prims_dev, factor)
            if r and not state["set"].valid():
                state["set"].close()
                state["set"], state["graph"] = record()
                recorded += 1
            with torch.cuda.stream(stream):
                state["graph"].replay()
            stream.synchronize()
            pms.append((time.perf_counter() - t0) * 1e3)
            rebuilt.append(bool(r))
            if check and f in (0, frames // 2, frames - 1):
                ok = True
                for i, (_, r_, tm) in enumerate(qb):
                    w = dbvh.raycast_nearest(r_, tm)
                    ctx.synchronize()
                    ok &= _same(state["set"].leaf[i].cpu(), w["leaf"].cpu()) and _same_bits(state["set"].toi[i].cpu(),
                                                                                            w["toi"].cpu())
                par.append(dict(frame=f, batches_equal_single_calls=bool(ok)))
        pm = float(np.median(pms))
        log(f"{frames} policy frames (factor {factor}): median {pm:.2f} ms, {sum(rebuilt)} rebuilds, "
            f"max {np.max(pms):.2f} ms, batches == single calls {all(p_['batches_equal_single_calls'] for p_ in par)}")
        return dict(factor=factor, frame_ms_median=round(pm, 3), frame_ms_first=round(pms[0], 3),
                    frame_ms_last=round(pms[-1], 3), frame_ms_max=round(float(np.max(pms)), 3),
                    rebuilds=int(sum(rebuilt)), rebuild_frames=[i for i, v in enumerate(rebuilt) if v],
                    graphs_recorded_again=recorded, mrays_per_s=round(float(sizes.sum()) / (pm * 1e-3) / 1e6, 1),
                    frame_ms=[round(v, 2) for v in pms], parity=par)

    runs = [policy_run(fc) for fc in (1.1, 1.25, 1.5, 2.0)]
    best = min(runs, key=lambda r_: r_["frame_ms_median"])
    out["policy_frames"] = best
    out["policy_factor_sweep"] = [dict(factor=r_["factor"], frame_ms_median=r_["frame_ms_median"],
                                       frame_ms_max=r_["frame_ms_max"], rebuilds=r_["rebuilds"]) for r_ in runs]
    state["set"].close()
    return out


def c4():
    import torch
    import sq_amd
    from sq_amd import workloads as W

    t0 = time.time()
    sc = W.gmm_boxes_scene()
    rays, regions, casts, lim = W.c4_queries(sc)
    log(f"generated in {time.time() - t0:.1f}s")
    out = dict(workload="c4: D=6, 1,000,000 GMM boxes; 1M rays, 262,144 overlaps, 262,144 box casts")
    ctx = sq_amd.Context(0)
    t0 = time.time()
    bvh = sq_amd.DeviceBvh.build(ctx, 6, sq_amd.AABB, sc.prims)
    torch.cuda.synchronize()
    out["build_s"] = round(time.time() - t0, 2)
    st = bvh.stats()
    out.update(tree_nodes=st["node_count"], tree_depth=st["max_depth"])
    r, q, cst, lm = _dev(rays), _dev(regions), _dev(casts), _dev(lim)

    def best_ms(fn, reps=5):
        fn()
        ctx.synchronize()
        ts = []
        for _ in range(reps):
            fn()
            ts.append(ctx.last_kernel_ms())
        return min(ts)

    res = {}
    ms = best_ms(lambda: res.__setitem__("ray", bvh.raycast_nearest(r, sc.max_toi)))
    out["nearest_ms"] = round(ms, 4)
    out["nearest_mrays"] = round(rays.shape[0] / (ms * 1e-3) / 1e6, 1)
    # overlap: count + fill + sort is several launches; time the whole call
    bvh.overlap(q)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        off, leaf = bvh.overlap(q)
    ctx.synchronize()
    oms = (time.perf_counter() - t0) * 1e3 / 3
    out.update(overlap_ms=round(oms, 3), overlap_mqueries=round(regions.shape[0] / (oms * 1e-3) / 1e6, 2),
               overlap_pairs=int(off[-1].item()))
    ms = best_ms(lambda: res.__setitem__("cast", bvh.shapecast_nearest(cst, 50.0, max_toi_per_cast=lm)))
    out["cast_ms"] = round(ms, 4)
    out["cast_mqueries"] = round(casts.shape[0] / (ms * 1e-3) / 1e6, 1)
    hr = res["ray"]["leaf"].cpu().numpy()
    hc = res["cast"]["leaf"].cpu().numpy()
    out.update(ray_hit_fraction=round(float((hr >= 0).mean()), 4), cast_hit_fraction=round(float((hc >= 0).mean()), 4))
    log(f"nearest {out['nearest_mrays']} Mrays/s, overlap {out['overlap_mqueries']} Mq/s ({out['overlap_pairs']} pairs), "
        f"casts {out['cast_mqueries']} Mq/s")
    return out


def build():
    """Device build (sq_bvh_build_device) against the host build + upload, and
    the device rebuild, on the scenes of the other workloads."""
    import torch
    import sq_amd
    from sq_amd import workloads as W

    ctx = sq_amd.Context(0)
    out = dict(workload="builds: C1 heightfield 2,097,152 triangles; C4 1,000,000 D=6 boxes; C2 100,000 moving boxes")
    scenes = [("c1_heightfield", 3, sq_amd.TRIANGLE, W.heightfield_scene().prims),
              ("c4_gmm_boxes_d6", 6, sq_amd.AABB, W.gmm_boxes_scene().prims),
              ("c2_moving_boxes", 3, sq_amd.AABB,
               W.boxes_scene(n=100_000, seed=7, lo=-1000.0, hi=1000.0, hmin=0.5, hmax=5.0).prims)]
    for name, dim, kind, prims in scenes:
        pd = _dev(prims)
        torch.cuda.synchronize()
        b = sq_amd.DeviceBvh.build_device(ctx, dim, kind, pd)      # warm-up (allocations, code load)
        ts = []
        for _ in range(5):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            b.rebuild(pd)
            ts.append((time.perf_counter() - t0) * 1e3)
        t0 = time.perf_counter()
        fresh = sq_amd.DeviceBvh.build_device(ctx, dim, kind, pd)
        t_fresh = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        hb = sq_amd.DeviceBvh.build(ctx, dim, kind, prims)
        t_host = (time.perf_counter() - t0) * 1e3
        st = fresh.stats()
        same = all(np.array_equal(x.view(np.uint32), y.view(np.uint32)) for x, y in zip(hb.export(), fresh.export()))
        out[name] = dict(prims=int(prims.shape[0]), nodes=st["node_count"], depth=st["max_depth"],
                         rebuild_ms=round(float(np.median(ts)), 2), device_build_ms=round(t_fresh, 2),
                         host_build_upload_ms=round(t_host, 1), speedup=round(t_host / float(np.median(ts)), 1),
                         same_tree_as_host=bool(same))
        log(f"{name}: rebuild {np.median(ts):.2f} ms, fresh device build {t_fresh:.2f} ms, host build + upload "
            f"{t_host:.1f} ms, same tree {same}")
        del b, fresh, hb
    return out


CONSUMER_BALL = r"""
__device__ bool sq_narrow(const float* p, const float* o, const float* d, float max_toi, float* toi) {
  float oc[SQ_RTC_D];
  for (int k = 0; k < SQ_RTC_D; ++k) oc[k] = o[k] - p[k];
  float b = 0.0f, cc = 0.0f;
  for (int k = 0; k < SQ_RTC_D; ++k) b = b + oc[k] * d[k];
  for (int k = 0; k < SQ_RTC_D; ++k) cc = cc + oc[k] * oc[k];
  cc = cc - p[SQ_RTC_D] * p[SQ_RTC_D];
  const float disc = b * b - cc;
  if (disc < 0.0f) return false;
  const float sq = __builtin_sqrtf(disc);
  float t = -b - sq;
  if (t < 0.0f) t = -b + sq;
  if (t < 0.0f || t > max_toi) return false;
  *toi = t;
  return true;
}
"""


def consumer():
    """A consumer narrow phase (the ball, compiled at run time) against the
    built-in SPHERE kind on the same scene: 1,000,000 balls, 4M rays."""
    import torch
    import sq_amd
    from sq_amd import workloads as W

    ctx = sq_amd.Context(0)
    s = W.Stream(0xC0C0)
    n = 1_000_000
    c = s.range(-500.0, 500.0, n * 3).reshape(n, 3)
    r = s.range(0.5, 3.0, n).reshape(n, 1)
    balls = np.concatenate([c, r], 1).astype(np.float32)
    boxes = np.concatenate([c - r, c + r], 1).astype(np.float32)
    rays = _dev(W.random_rays(4 << 20, 0xC0C1, -520.0, 520.0))
    t0 = time.perf_counter()
    nw = sq_amd.Narrow(ctx, 3, CONSUMER_BALL, 4)
    compile_s = time.perf_counter() - t0
    cb = nw.build(boxes, balls)
    bb = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.SPHERE, balls)

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            out = fn()
            e1.record()
            e1.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts)), out

    ms_c, oc = timed(lambda: cb.raycast_nearest(rays, 2000.0))
    ms_b, ob = timed(lambda: bb.raycast_nearest(rays, 2000.0))
    same = bool(torch.equal(oc["leaf"], ob["leaf"]) and torch.equal(oc["toi"].view(torch.int32), ob["toi"].view(torch.int32)))
    out = dict(workload="consumer narrow phase: 1,000,000 balls, 4,194,304 rays, nearest-hit",
               compile_s=round(compile_s, 2), consumer_ms=round(ms_c, 3), consumer_mrays=round(rays.shape[0] / ms_c / 1e3, 1),
               builtin_ms=round(ms_b, 3), builtin_mrays=round(rays.shape[0] / ms_b / 1e3, 1), same_results=same,
               hit_fraction=round(float((ob["leaf"] >= 0).float().mean()), 4))
    log(f"consumer {ms_c:.2f} ms vs built-in {ms_b:.2f} ms, same {same}, compile {compile_s:.2f} s")
    return out


def dispatch():
    """The HIP backend's constants for the reference's dispatch cost model
    (crates/spatial-query/src/dispatch.rs:38-116): the host round trip of a
    nearest-hit batch through sq_amd.dispatch.HipBackend (rays through pinned
    host memory, leaf + toi back, synchronised) on a 100,000-ball scene.  A
    batch of up to ~100 rays takes as long as its longest ray's chain of
    dependent node fetches, so the constants are the line through the round
    trips of 256 and 1M rays: GPU_FIXED_US its intercept, the per-ray slope over
    depth^2 (depth = log2(leaves + 1)) split between traversal and narrow test
    in the reference's proportion."""
    import torch  # noqa: F401
    import sq_amd
    from sq_amd import dispatch as dp
    from sq_amd import workloads as W

    ctx = sq_amd.Context(0)
    s = W.Stream(0xD15)
    n = 100_000
    c = s.range(-100.0, 100.0, n * 3).reshape(n, 3)
    r = s.range(0.1, 2.0, n).reshape(n, 1)
    bvh = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.SPHERE, np.concatenate([c, r], 1).astype(np.float32))
    big = 1 << 20
    rays = W.random_rays(big, 0xD16, -120.0, 120.0)
    be = dp.HipBackend(bvh)

    res = {}
    for b in (1, 4, 16, 64, 256, 1024, 4096, 65536, big):
        reps = 50 if b <= 4096 else 10
        be.raycast_nearest_batch(rays[:b], 400.0)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            be.raycast_nearest_batch(rays[:b], 400.0)
            ts.append((time.perf_counter() - t0) * 1e6)
        res[b] = float(np.median(ts))
    depth = dp.depth(n)
    per_ray = (res[big] - res[256]) / (big - 256)
    fixed = res[256] - 256 * per_ray
    costs = dp.fit_gpu(dp.Costs.reference(), fixed, per_ray, n)
    d = dp.Dispatcher(dp.Costs.reference())
    d.register(dp.HipBackend(bvh, costs))
    crossover = d.crossover(n)
    out = dict(workload="dispatch cost constants: 100,000 balls (depth log2(leaves+1) = %.2f), host round trip "
                        "through HipBackend (pinned rays in, leaf + toi out, synchronised)" % depth,
               round_trip_us={str(k): round(v, 1) for k, v in res.items()},
               GPU_FIXED_US=round(fixed, 1), per_ray_us=round(per_ray, 6),
               GPU_TRAVERSAL_US_PER_LEVEL2=float("%.3g" % costs.gpu_traversal_us_per_level2),
               GPU_NARROW_US_PER_LEVEL2=float("%.3g" % costs.gpu_narrow_us_per_level2),
               fit="line through the round trips of 256 and 1,048,576 rays",
               reference_GPU_FIXED_US=1350.0, reference_per_ray_us=0.053,
               model_crossover_batch_vs_reference_cpu=crossover)
    log(f"fixed {fixed:.1f} us, per ray {per_ray * 1e3:.3f} ns, model crossover vs the reference CPU constants "
        f"{crossover} rays")
    return out


def f64():
    """The `f64` feature against FP32 on the same scenes: 1,000,000 balls in
    D = 3 (centres U[-1000,1000]^3, radii U[0.5,4]) with 4M rays, and the C4 D = 6
    box scene with its 1M rays; nearest and any-hit rays, plus all hits and
    overlap on the ball scene.  Kernel times from events on the query stream."""
    import torch
    import sq_amd
    from sq_amd import workloads as W

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
        return best

    ctx = sq_amd.Context(0)
    out = dict(workload="f64 vs f32: 1M balls D=3 with 4M rays; C4 (1M GMM boxes D=6) with 1M rays")
    rng = np.random.default_rng(0xF64B)
    n, nr = 1_000_000, 1 << 22
    balls = np.concatenate([rng.uniform(-1000, 1000, (n, 3)), rng.uniform(0.5, 4.0, (n, 1))], 1)
    o = rng.uniform(-1000, 1000, (nr, 3))
    d = rng.normal(size=(nr, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1)
    scenes = [("balls", 3, sq_amd.SPHERE, balls, rays, 4000.0)]
    sc = W.gmm_boxes_scene()
    c4_rays, _, _, _ = W.c4_queries(sc)
    scenes.append(("c4_boxes", 6, sq_amd.AABB, sc.prims.astype(np.float64), c4_rays.astype(np.float64), sc.max_toi))
    for name, dim, kind, prims, q, max_toi in scenes:
        t0 = time.time()
        b64 = sq_amd.DeviceBvh64(ctx, dim, kind, prims)
        build64 = time.time() - t0
        b32 = sq_amd.DeviceBvh.build(ctx, dim, kind, prims.astype(np.float32))
        q64 = torch.from_numpy(np.ascontiguousarray(q)).cuda()
        q32 = q64.float().contiguous()
        res = {}
        ms64 = timed(lambda: res.__setitem__("f64", b64.raycast_nearest(q64, max_toi)))
        ms32 = timed(lambda: res.__setitem__("f32", b32.raycast_nearest(q32, max_toi)))
        any64 = timed(lambda: b64.raycast_any(q64, max_toi))
        any32 = timed(lambda: b32.raycast_any(q32, max_toi))
        l64, l32 = res["f64"]["leaf"].cpu().numpy(), res["f32"]["leaf"].cpu().numpy()
        m = q.shape[0] / 1e6
        out[name] = dict(prims=prims.shape[0], rays=q.shape[0], build64_s=round(build64, 2),
                         nearest_f64_mrays=round(m / (ms64 * 1e-3), 1), nearest_f32_mrays=round(m / (ms32 * 1e-3), 1),
                         any_f64_mrays=round(m / (any64 * 1e-3), 1), any_f32_mrays=round(m / (any32 * 1e-3), 1),
                         f64_over_f32=round(ms32 / ms64, 3), hit_fraction=round(float((l64 >= 0).mean()), 4),
                         same_leaf_as_f32=round(float((l64 == l32).mean()), 6))
        if name == "balls":
            sub = q64[: 1 << 20].contiguous()
            all_ms = timed(lambda: b64.raycast_all(sub, max_toi), reps=3)
            off, _, _ = b64.raycast_all(sub, max_toi)
            h = rng.uniform(2.0, 20.0, (1 << 18, 3))
            c = rng.uniform(-1000, 1000, (1 << 18, 3))
            reg = torch.from_numpy(np.concatenate([c - h, c + h], 1)).cuda()
            ov_ms = timed(lambda: b64.overlap(reg), reps=3)
            off2, _ = b64.overlap(reg)
            out[name].update(all_hits_1m_rays_ms=round(all_ms, 3), all_hits_entries=int(off[-1].item()),
                             overlap_262k_ms=round(ov_ms, 3), overlap_pairs=int(off2[-1].item()))
        log(f"{name}: f64 {out[name]['nearest_f64_mrays']} Mrays/s, f32 {out[name]['nearest_f32_mrays']} Mrays/s")
        b64.close()
        b32.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["c2", "c4", "build", "consumer", "dispatch", "f64"])
    ap.add_argument("--frames", type=int, default=60)
    ap.add_argument("--batches", type=int, default=256)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    a = ap.parse_args()
    if a.which == "c2":
        res = c2(frames=a.frames, n_batches=a.batches, tune=not a.no_tune, check=not a.no_check)
    elif a.which == "c4":
        res = c4()
    elif a.which == "build":
        res = build()
    elif a.which == "consumer":
        res = consumer()
    elif a.which == "f64":
        res = f64()
    else:
        res = dispatch()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
