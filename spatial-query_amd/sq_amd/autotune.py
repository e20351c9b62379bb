"""Autotuner for the engine's kernel tuning: block size, rays per lane (pool
size), LDS ring-stack depth, Morton ray order and the packet walk (one
wave-coherent walk per 64 rays, for coherent batches such as camera rays).

Results never depend on the tuning, only speed does; the tuner checks that too:
every candidate's results must equal the first candidate's bit for bit.

    best = autotune(ctx, prepare)          # leaves ctx set to the fastest

`prepare()` is called once per candidate after the context has been set to it
(not timed: build batch sets there, since they read the tuning at creation) and
returns `go`, a callable that enqueues the workload once and returns the result
tensors to compare.
"""
from __future__ import annotations

import itertools
import time

DEFAULT_SPACE = dict(block=(64, 128, 256), rays_per_lane=(4, 8, 16), lds_stack=(4, 8, 16), ray_order=(0, 1), packet=(0, 1))


def _same(a, b) -> bool:
    import torch
    if len(a) != len(b):
        return False
    for x, y in zip(a, b):
        if x.shape != y.shape or x.dtype != y.dtype:
            return False
        if x.dtype.is_floating_point:
            x = x.view(torch.int32) if x.dtype == torch.float32 else x
            y = y.view(torch.int32) if y.dtype == torch.float32 else y
        if not torch.equal(x, y):
            return False
    return True


def autotune(ctx, prepare, space: dict | None = None, reps: int = 3, log=None) -> dict:
    """Time every combination of `space` (defaults: DEFAULT_SPACE) on the
    workload and set the context to the fastest.  Returns dict(best=tuning,
    ms=its time, table=[(tuning, ms)], baseline=ms of the context's tuning on
    entry)."""
    import torch
    space = dict(DEFAULT_SPACE if space is None else space)
    entry = ctx.tuning()
    names = list(space)
    cands = [dict(entry)] + [dict(entry, **dict(zip(names, vals))) for vals in itertools.product(*(space[k] for k in names))]
    seen, table, ref = set(), [], None
    for cfg in cands:
        key = tuple(sorted(cfg.items()))
        if key in seen:
            continue
        seen.add(key)
        ctx.set_tuning(**cfg)
        go = prepare()
        out = go()                      # warm-up, sizes every buffer
        ctx.synchronize()
        torch.cuda.synchronize()
        got = tuple(t.clone() for t in out)
        if ref is None:
            ref = got
        elif not _same(ref, got):
            ctx.set_tuning(**entry)
            raise AssertionError(f"tuning {cfg} changed the results")
        t0 = time.perf_counter()
        for _ in range(reps):
            go()
        ctx.synchronize()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        table.append((cfg, ms))
        if log:
            log(f"tuning {cfg}: {ms:.3f} ms")
    best_cfg, best_ms = min(table, key=lambda r: r[1])
    ctx.set_tuning(**best_cfg)
    return dict(best=best_cfg, ms=best_ms, table=table, baseline=table[0][1])
