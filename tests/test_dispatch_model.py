"""The dispatcher and cost model (sq_amd.dispatch) against the reference's own
tests (crates/spatial-query/tests/dispatch.rs), with the reference's constants
(this pins the restatement: a wrong depth, a dropped hint or a swapped
guardrail fails one of them), then with this engine's MI355X GPU constants.
Host logic only: no GPU."""
import json
import math
import os

import pytest

from sq_amd import dispatch as dp
from sq_amd.dispatch import BackendKind as K, Costs, QueryContext as Q

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def full(costs):
    """tests/dispatch.rs:172-177: CPU, then the draw-pass and GPU stubs."""
    d = dp.Dispatcher(costs)
    d.register(dp.CostBackend(K.DRAW_PASS, costs))
    d.register(dp.CostBackend(K.GPU, costs))
    return d


def test_depth_and_cost_shapes():
    c = Costs.reference()
    assert dp.depth(0) == 1.0 and dp.depth(1) == 1.0 and dp.depth(3) == 2.0
    assert dp.depth(100_000) == pytest.approx(16.6096, abs=1e-4)
    # the sphere calibration the constants were fitted to (dispatch.rs:46-78):
    # 2.70 us/ray on the CPU and 0.053 us/ray on the GPU at depth ~16.6
    ctx = Q(1, 1, 100_000)
    assert dp.cpu_batch_us(c, ctx) == pytest.approx(2.70, rel=0.01)
    assert dp.gpu_batch_us(c, ctx) - c.gpu_fixed_us == pytest.approx(0.053, rel=0.01)
    assert dp.cpu_batch_us(c, Q(1000, 1000, 100_000)) == pytest.approx(1000 * dp.cpu_batch_us(c, ctx))
    # the hint scales the narrow part only
    d2 = dp.depth(1000) ** 2
    assert dp.cpu_batch_us(c, Q(1, 1, 1000).with_narrow_cost(3.0)) == pytest.approx((0.0020 + 3 * 0.0078) * d2)
    assert dp.draw_pass_batch_us(c, Q(10, 1, 5000)) == 0.5
    assert dp.draw_pass_batch_us(c, Q(10, 4, 5000)) == pytest.approx(0.5 + 3 * 0.002 * 5000)
    assert Q(1, 1, 1).with_narrow_cost(-2.0).narrow_cost_hint == 0.0


def test_reference_routing_tests():
    """deterministic_request_pins_cpu, no_device_falls_back_to_cpu,
    single_view_large_scene_picks_draw_pass, scattered_large_batch_picks_gpu,
    small_batch_stays_cpu_without_alternates_winning, force_override_wins_over_everything."""
    c = Costs.reference()
    d = full(c)
    assert d.select(Q(100_000, 100_000, 100_000).with_device(True).deterministic(True)) is K.CPU
    assert d.select(Q(1, 1, 100_000).with_device(False)) is K.CPU
    assert d.select(Q(1, 1, 100_000).with_device(True)) is K.DRAW_PASS
    assert d.select(Q(5_000, 1, 100_000).with_device(True)) is K.DRAW_PASS
    assert d.select(Q(100_000, 100_000, 100_000).with_device(True)) is K.GPU
    assert dp.Dispatcher(c).select(Q(1, 1, 100_000).with_device(True)) is K.CPU
    d.force(K.GPU)
    assert d.select(Q(1, 1, 10).deterministic(True)) is K.GPU
    d.force(None)
    assert d.select(Q(1, 1, 10).deterministic(True)) is K.CPU


def test_reference_crossover_band_and_field_leaves():
    """calibrated_model_crosses_over_in_the_measured_band and
    expensive_field_leaves_route_to_gpu_sooner with the reference's constants."""
    d = full(Costs.reference())
    sc = lambda b, leaves: Q(b, b, leaves).with_device(True)
    assert d.select(sc(64, 100_000)) is K.CPU
    assert d.select(sc(8_000, 100_000)) is K.GPU
    x = d.crossover(100_000)
    assert 256 <= x <= 2048
    assert x == 512
    cheap = Q(256, 256, 100_000).with_device(True)
    assert d.select(cheap) is K.CPU
    assert d.select(cheap.with_narrow_cost(30.0)) is K.GPU
    assert d.crossover(100_000, hint=30.0) < x


def test_mi355x_constants_move_the_crossover():
    """With this engine's measured GPU constants (and the reference's CPU
    constants) the model routes a scattered 100k-leaf batch to the GPU from 256
    rays: the measured fixed cost is 3.8x lower than the reference's.  The
    constants agree with the committed measurement, and refitting them from
    its figures gives the same constants back."""
    with open(os.path.join(ROOT, "profiles", "dispatch_cost_constants.json")) as fh:
        m = json.load(fh)
    c = Costs.mi355x()
    assert c.gpu_fixed_us == m["GPU_FIXED_US"]
    assert c.gpu_traversal_us_per_level2 == m["GPU_TRAVERSAL_US_PER_LEVEL2"]
    assert c.gpu_narrow_us_per_level2 == m["GPU_NARROW_US_PER_LEVEL2"]
    f = dp.fit_gpu(Costs.reference(), m["GPU_FIXED_US"], m["per_ray_us"], 100_000)
    assert f.gpu_traversal_us_per_level2 == pytest.approx(c.gpu_traversal_us_per_level2, rel=5e-3)
    assert f.gpu_narrow_us_per_level2 == pytest.approx(c.gpu_narrow_us_per_level2, rel=5e-3)
    rt = {int(k): v for k, v in m["round_trip_us"].items()}
    slope = (rt[1 << 20] - rt[256]) / ((1 << 20) - 256)
    assert slope == pytest.approx(m["per_ray_us"], rel=1e-3)
    assert rt[256] - 256 * slope == pytest.approx(c.gpu_fixed_us, abs=0.1)
    d = full(c)
    x = d.crossover(100_000)
    assert x == 256
    assert x == m["model_crossover_batch_vs_reference_cpu"]
    assert d.select(Q(128, 128, 100_000).with_device(True)) is K.CPU
    # a HipBackend registers as the Gpu kind with these constants
    h = dp.HipBackend(bvh=None)
    assert h.kind is K.GPU and h.costs == c
    assert not h.can_serve(Q(10, 10, 10)) and h.can_serve(Q(10, 10, 10).with_device(True))


def test_fit_cpu_round_trips_the_reference_constants():
    c = Costs.reference()
    per_ray = dp.cpu_batch_us(c, Q(1, 1, 100_000))
    f = dp.fit_cpu(Costs.mi355x(), per_ray, 100_000)
    assert f.cpu_traversal_us_per_level2 == pytest.approx(0.0020)
    assert f.cpu_narrow_us_per_level2 == pytest.approx(0.0078)
    assert math.isclose(dp.cpu_batch_us(f, Q(7, 7, 100_000)), 7 * per_ray)
