"""The reference's crossover test (tests/dispatch.rs:228-255,
calibrated_model_crosses_over_in_the_measured_band) restated on MI355X: on a
100,000-ball scene, the host round trip of this engine's GPU backend
(HipBackend: host rays in, host results out) and the sequential CPU traversal
(the oracle's restatement of the reference's raycast_nearest, one thread) are
measured on the same machine; the cost model, refitted from those two
measurements (the GPU line through the round trips of 256 rays, where the
longest ray's dependent fetches rather than the batch size set the time, and
of 16,384 rays, as sq_amd.benches dispatch fits it; the slopes over depth^2
split in the reference's proportion), must route to the GPU from a batch in
the band around the measured crossover: at most 2x below it and at most 4x
above it, the reference's own band (256..2048 around its measured ~510)."""
import time

import numpy as np
import pytest

import oracle
import sq_amd
from sq_amd import dispatch as dp
from sq_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_calibrated_model_crosses_over_in_the_measured_band(ctx):
    print("case: dispatch crossover, 100,000 balls", flush=True)
    s = W.Stream(0xD15)
    n = 100_000
    c = s.range(-100.0, 100.0, n * 3).reshape(n, 3)
    r = s.range(0.1, 2.0, n).reshape(n, 1)
    balls = np.concatenate([c, r], 1).astype(np.float32)
    bvh = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.SPHERE, balls)
    gpu = dp.HipBackend(bvh)
    rays = W.random_rays(1 << 15, 0xD16, -120.0, 120.0)
    o = oracle.Scene(3, oracle.SPHERE, balls)
    # the backend returns the reference's results
    leaf, toi = gpu.raycast_nearest_batch(rays[:4096], 400.0)
    ref = o.raycast_nearest(rays[:4096], 400.0, threads=8)
    assert (ref["leaf"] >= 0).sum() > 100
    assert np.array_equal(leaf.astype(np.int64), ref["leaf"])
    assert np.array_equal(toi.view(np.uint32), ref["t"].view(np.uint32))

    def median_us(fn, reps):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append((time.perf_counter() - t0) * 1e6)
        return float(np.median(ts))

    sizes = [1 << k for k in range(0, 15)]
    gpu_us = {b: median_us(lambda b=b: gpu.raycast_nearest_batch(rays[:b], 400.0), 30) for b in sizes}
    cpu_n = 8192
    cpu_per_ray = median_us(lambda: o.raycast_nearest(rays[:cpu_n], 400.0, threads=1), 5) / cpu_n
    measured = next((b for b in sizes[1:] if gpu_us[b] < cpu_per_ray * b), None)
    assert measured is not None, f"the GPU never won: {gpu_us}, cpu {cpu_per_ray:.3f} us/ray"
    slope = (gpu_us[sizes[-1]] - gpu_us[256]) / (sizes[-1] - 256)
    costs = dp.fit_cpu(dp.fit_gpu(dp.Costs.reference(), gpu_us[256] - 256 * slope, slope, n), cpu_per_ray, n)
    d = dp.Dispatcher(costs)
    d.register(dp.HipBackend(bvh, costs))
    model = d.crossover(n)
    print(f"round trip us {dict((b, round(v, 1)) for b, v in gpu_us.items())}; cpu {cpu_per_ray:.3f} us/ray; "
          f"measured crossover {measured}, model crossover {model}; fitted {costs}", flush=True)
    assert model is not None
    assert measured / 2 <= model <= measured * 4
    # a tiny batch stays on the CPU, a large one goes to the GPU, as in the reference's test
    assert d.select(dp.QueryContext(1, 1, n).with_device(True)) is dp.BackendKind.CPU
    assert d.select(dp.QueryContext(8000, 8000, n).with_device(True)) is dp.BackendKind.GPU
    bvh.close()
