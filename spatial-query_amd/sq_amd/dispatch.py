"""The reference's backend cost model and dispatcher (crates/spatial-query/src/
dispatch.rs), with this engine registered as its `Gpu` backend.

The reference routes a batch of nearest-hit rays to the cheapest registered
`QueryBackend` by `estimate_cost` (microseconds), after two guardrails: a
forced backend wins, and a deterministic request pins the CPU
(dispatch.rs:400-414).  Its cost functions share one shape,
`fixed + batch * (traversal + hint * narrow) * depth^2` with
`depth = max(1, log2(leaves + 1))` (dispatch.rs:38-116).  `Costs` holds the
constants: `Costs.reference()` the reference's (fitted on its own machine),
`Costs.mi355x()` the same CPU constants with the GPU constants this engine
measured on MI355X (`python -m sq_amd.benches dispatch`: HipBackend's host
round trip on a 100,000-ball scene, the line through batches of 256 and 1M
rays, since a batch of up to ~100 rays takes as long as its longest ray's
chain of dependent node fetches).  `fit_cpu` / `fit_gpu` refit either side
from measured times, splitting each per-ray slope between traversal and narrow
test in the reference's proportion.

`HipBackend` is the `Gpu` backend a Rust binding registers: `can_serve` when a
device is available, `estimate_cost` with the MI355X constants, and
`raycast_nearest_batch` runs host rays through the engine and returns host
results.  The CPU backend is the reference's own traversal; here it is a cost
only (`CpuCost`).
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass, field, replace


class BackendKind(enum.Enum):
    """dispatch.rs:128-141."""
    CPU = "cpu"
    DRAW_PASS = "draw_pass"
    GPU = "gpu"


@dataclass(frozen=True)
class QueryContext:
    """dispatch.rs:143-190: the batch the dispatcher routes."""
    batch_size: int
    distinct_directions: int
    scene_leaves: int
    device_available: bool = False
    require_deterministic: bool = False
    narrow_cost_hint: float = 1.0

    def with_device(self, available: bool) -> "QueryContext":
        return replace(self, device_available=bool(available))

    def with_narrow_cost(self, hint: float) -> "QueryContext":
        return replace(self, narrow_cost_hint=max(float(hint), 0.0))

    def deterministic(self, required: bool) -> "QueryContext":
        return replace(self, require_deterministic=bool(required))


@dataclass(frozen=True)
class Costs:
    """The cost model's constants in microseconds (dispatch.rs:38-83)."""
    cpu_traversal_us_per_level2: float
    cpu_narrow_us_per_level2: float
    gpu_fixed_us: float
    gpu_traversal_us_per_level2: float
    gpu_narrow_us_per_level2: float
    draw_pass_readback_us: float = 0.5
    draw_pass_us_per_leaf: float = 0.002

    @staticmethod
    def reference() -> "Costs":
        return Costs(0.0020, 0.0078, 1350.0, 0.00004, 0.000153)

    @staticmethod
    def mi355x() -> "Costs":
        """The reference's CPU constants with this engine's GPU constants on
        MI355X (100,000 balls, depth 16.61, HipBackend's host round trip: the
        line through 256 rays, 351.3 us, and 1,048,576 rays, 2,620.7 us; a
        batch of one takes 63.9 us)."""
        return replace(Costs.reference(), gpu_fixed_us=350.8, gpu_traversal_us_per_level2=1.63e-06,
                       gpu_narrow_us_per_level2=6.22e-06)


def depth(leaves: int) -> float:
    """dispatch.rs:85-89: log2(leaves + 1), at least 1."""
    return max(math.log2(leaves + 1), 1.0)


def cpu_batch_us(c: Costs, ctx: QueryContext) -> float:
    d = depth(ctx.scene_leaves)
    return ctx.batch_size * (c.cpu_traversal_us_per_level2 + ctx.narrow_cost_hint * c.cpu_narrow_us_per_level2) * d * d


def gpu_batch_us(c: Costs, ctx: QueryContext) -> float:
    d = depth(ctx.scene_leaves)
    return c.gpu_fixed_us + ctx.batch_size * (c.gpu_traversal_us_per_level2 + ctx.narrow_cost_hint *
                                              c.gpu_narrow_us_per_level2) * d * d


def draw_pass_batch_us(c: Costs, ctx: QueryContext) -> float:
    extra = max(ctx.distinct_directions - 1, 0)
    return c.draw_pass_readback_us + extra * c.draw_pass_us_per_leaf * ctx.scene_leaves


def _split(coef: float, trav: float, narrow: float) -> tuple[float, float]:
    return coef * trav / (trav + narrow), coef * narrow / (trav + narrow)


def fit_cpu(c: Costs, per_ray_us: float, leaves: int) -> Costs:
    """CPU constants from a measured sequential per-ray time at hint 1."""
    t, n = _split(per_ray_us / depth(leaves) ** 2, c.cpu_traversal_us_per_level2, c.cpu_narrow_us_per_level2)
    return replace(c, cpu_traversal_us_per_level2=t, cpu_narrow_us_per_level2=n)


def fit_gpu(c: Costs, fixed_us: float, per_ray_us: float, leaves: int) -> Costs:
    """GPU constants from a measured batch-of-one time and per-ray slope at hint 1."""
    t, n = _split(per_ray_us / depth(leaves) ** 2, c.gpu_traversal_us_per_level2, c.gpu_narrow_us_per_level2)
    return replace(c, gpu_fixed_us=fixed_us, gpu_traversal_us_per_level2=t, gpu_narrow_us_per_level2=n)


@dataclass
class CostBackend:
    """A backend as the dispatcher weighs it: kind, whether it can serve a
    context, and its cost.  The CPU backend always serves (dispatch.rs:255-269);
    the draw-pass and GPU backends need a device."""
    kind: BackendKind
    costs: Costs

    def can_serve(self, ctx: QueryContext) -> bool:
        return True if self.kind is BackendKind.CPU else ctx.device_available

    def estimate_cost(self, ctx: QueryContext) -> float:
        if self.kind is BackendKind.CPU:
            return cpu_batch_us(self.costs, ctx)
        if self.kind is BackendKind.GPU:
            return gpu_batch_us(self.costs, ctx)
        return draw_pass_batch_us(self.costs, ctx)


def CpuCost(costs: Costs) -> CostBackend:
    return CostBackend(BackendKind.CPU, costs)


@dataclass
class Dispatcher:
    """dispatch.rs:351-414: the CPU backend is always registered; select()
    returns the forced kind if any, CPU for a deterministic request, else the
    cheapest backend that can serve (ties to the earlier registered, as Rust's
    min_by keeps the first minimum), CPU when none can."""
    costs: Costs
    backends: list = field(default_factory=list)
    forced: BackendKind | None = None

    def __post_init__(self):
        if not self.backends:
            self.backends = [CpuCost(self.costs)]

    def register(self, backend) -> None:
        self.backends.append(backend)

    def force(self, kind: BackendKind | None) -> None:
        self.forced = kind

    def select(self, ctx: QueryContext) -> BackendKind:
        if self.forced is not None:
            return self.forced
        if ctx.require_deterministic:
            return BackendKind.CPU
        best, best_cost = None, None
        for b in self.backends:
            if not b.can_serve(ctx):
                continue
            c = b.estimate_cost(ctx)
            if best is None or c < best_cost:
                best, best_cost = b, c
        return best.kind if best is not None else BackendKind.CPU

    def crossover(self, leaves: int, hint: float = 1.0, limit: int = 1 << 30) -> int | None:
        """The smallest power-of-two scattered batch (distinct directions =
        batch, device available) routed to the GPU, as the reference's
        crossover test searches it (tests/dispatch.rs:240-246)."""
        b = 2
        while b <= limit:
            ctx = QueryContext(b, b, leaves).with_device(True).with_narrow_cost(hint)
            if self.select(ctx) is BackendKind.GPU:
                return b
            b <<= 1
        return None


class HipBackend(CostBackend):
    """This engine as the reference's `Gpu` QueryBackend over one DeviceBvh:
    cost from the MI355X constants, and a host-in, host-out nearest-hit batch
    (the round trip the constants are measured on): the rays go through a
    pinned staging buffer to the device, the walk runs, leaf and toi come back
    through pinned buffers, all on torch's current stream (the context must
    enqueue there, the Context default), then it synchronises.  The buffers
    grow to the largest batch seen and are reused."""

    def __init__(self, bvh, costs: Costs | None = None):
        super().__init__(BackendKind.GPU, costs or Costs.mi355x())
        self.bvh = bvh
        self._cap = 0

    def _ensure(self, n: int) -> None:
        if n <= self._cap:
            return
        import torch
        cap = max(n, 1024)
        w = 2 * self.bvh.dim
        self._rays_h = torch.empty((cap, w), dtype=torch.float32).pin_memory()
        self._leaf_h = torch.empty(cap, dtype=torch.int32).pin_memory()
        self._toi_h = torch.empty(cap, dtype=torch.float32).pin_memory()
        self._rays_d = torch.empty((cap, w), dtype=torch.float32, device="cuda")
        self._leaf_d = torch.empty(cap, dtype=torch.int32, device="cuda")
        self._toi_d = torch.empty(cap, dtype=torch.float32, device="cuda")
        self._cap = cap

    def raycast_nearest_batch(self, rays, max_toi: float):
        """rays: host float32 [n][2D]; returns (leaf int32 [n], toi float32 [n]) on the host, -1 / inf on a miss."""
        import numpy as np
        import torch
        n = int(rays.shape[0])
        self._ensure(n)
        stream = torch.cuda.current_stream()
        if self.bvh.ctx.stream_ptr() != stream.cuda_stream:
            raise ValueError("HipBackend: the context must enqueue on torch's current stream")
        self._rays_h.numpy()[:n] = np.asarray(rays, np.float32).reshape(n, -1)
        self._rays_d[:n].copy_(self._rays_h[:n], non_blocking=True)
        self.bvh.raycast_nearest_into(self._rays_d[:n], float(max_toi), self._leaf_d[:n], self._toi_d[:n])
        self._leaf_h[:n].copy_(self._leaf_d[:n], non_blocking=True)
        self._toi_h[:n].copy_(self._toi_d[:n], non_blocking=True)
        stream.synchronize()
        return self._leaf_h.numpy()[:n].copy(), self._toi_h.numpy()[:n].copy()
