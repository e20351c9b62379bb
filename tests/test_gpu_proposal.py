"""Learned subtree proposal on the GPU, through the C ABI.

  * the matrix-core GEMM against numpy on the same bf16 inputs (a wrong MFMA
    lane layout, a dropped K block or a split-K slip fails it);
  * the network's loss and gradients against the MLP oracle (same bf16
    rounding points; tolerance 1e-2 relative per tensor, written below);
  * fused top-4 inference against the oracle's logits;
  * the proposal-seeded walk bit-exact against the plain walk and the CPU
    oracle whatever the network predicts (untrained, random ids, empty slots),
    under SQ_HITS_GUARDED, the mode whose answer does not depend on the order
    in which subtrees are walked (this module runs its walks in it);
  * what a trained network gains over an untrained one.
"""
import numpy as np
import pytest

import oracle
import sq_amd
from oracle import mlp
from sq_amd import workloads as W

from helpers import assert_nearest_equal

pytestmark = pytest.mark.gpu


def dev(a, dt=np.float32):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dt)).cuda()


def rel(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.fixture(scope="module", autouse=True)
def guarded_mode(ctx):
    ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED)
    yield
    ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)


@pytest.fixture(scope="module")
def small(ctx):
    sc = W.heightfield_scene(grid=257, extent=128.0)
    bvh = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, sc.prims)
    yield sc, bvh
    bvh.close()


@pytest.mark.parametrize("M,N,K,splits", [(128, 128, 32, 1), (256, 384, 160, 1), (384, 128, 1024, 4), (128, 256, 4096, 7)])
def test_mfma_gemm_matches_numpy(ctx, M, N, K, splits):
    print(f"case: mfma gemm {M}x{N}x{K} splits {splits}", flush=True)
    rng = np.random.default_rng(M * 7 + N * 3 + K + splits)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((N, K)).astype(np.float32)
    C = sq_amd.mfma_gemm(ctx, dev(A), dev(B), splits).cpu().numpy()
    Ab = mlp.bf16_round(A).astype(np.float64)
    Bb = mlp.bf16_round(B).astype(np.float64)
    ref = Ab @ Bb.T
    # bf16 x bf16 products are exact in fp32; K fp32 additions err by at most
    # K * 2^-24 * sum |a b|
    tol = K * 2.0 ** -24 * (np.abs(Ab) @ np.abs(Bb).T) + 1e-30
    assert np.all(np.abs(C - ref) <= tol)


def _labelled(bvh, p, rays):
    r = dev(rays)
    hit = bvh.raycast_nearest(r, 2000.0)
    return r, p.labels(hit["leaf"]), hit


@pytest.mark.parametrize("cut", [5, 9])
def test_gradients_match_oracle(ctx, small, cut):
    print(f"case: proposal gradients, cut depth {cut}", flush=True)
    sc, bvh = small
    p = sq_amd.Proposal(bvh, cut_depth=cut, seed=0xAB + cut)
    rays = W.proposal_training_rays(sc, 1000, seed=0x6AD + cut)       # 1000: padded to 1024 rows
    r, lab, _ = _labelled(bvh, p, rays)
    lab_h = lab.cpu().numpy()
    assert (lab_h >= 0).sum() > 200 and (lab_h < 0).sum() > 0
    layers = p.weights()
    B, box = p.features_spec()
    F = mlp.features(rays, B, box)
    loss_ref, ref = mlp.loss_and_grads(F, layers, lab_h)
    got = p.gradients(r, lab)
    # Same bf16 rounding points on both sides; the fp32 accumulation order
    # differs, which flips the odd bf16 rounding: 1e-2 relative per tensor.
    for l, ((gw, gb), (rw, rb)) in enumerate(zip(got, ref)):
        assert gw.shape == rw.shape and gb.shape == rb.shape
        assert rel(gw, rw) <= 1e-2, f"dW{l + 1}: {rel(gw, rw)}"
        assert rel(gb, rb) <= 1e-2, f"db{l + 1}: {rel(gb, rb)}"
    loss = p.train_step(r, lab, lr=0.0, want_loss=True)      # lr 0: weights unchanged
    assert loss == pytest.approx(loss_ref, rel=1e-3)
    p.close()


def test_adam_step_matches_oracle(ctx, small):
    """One training step from fresh moments equals Adam (lr, 0.9, 0.999, 1e-8)
    applied in float64 to the oracle's gradients: every weight moves by
    lr * sign(g) (m/c1 = g, v/c2 = g^2), within the fp32 update rounding."""
    print("case: one Adam step", flush=True)
    sc, bvh = small
    p = sq_amd.Proposal(bvh, cut_depth=6, seed=0xADA)
    rays = W.proposal_training_rays(sc, 4096, seed=0xADA)
    r, lab, _ = _labelled(bvh, p, rays)
    w0 = p.weights()
    B, box = p.features_spec()
    _, g = mlp.loss_and_grads(mlp.features(rays, B, box), w0, lab.cpu().numpy())
    lr = 1e-3
    p.train_step(r, lab, lr)
    w1 = p.weights()
    for l in range(4):
        for k in range(2):
            gg = np.asarray(g[l][k], np.float64)
            m, v = 0.1 * gg, 0.001 * gg * gg
            want = w0[l][k] - lr * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
            # where |g| is within a bf16 rounding of 0 the sign can differ: compare the rest
            big = np.abs(gg) > 1e-2 * np.abs(gg).max()
            assert np.abs(w1[l][k] - want)[big].max() <= 1e-6, (l, k)
    p.close()


@pytest.mark.parametrize("cut", [1, 9])
def test_infer_top4_matches_oracle(ctx, small, cut):
    print(f"case: fused top-4 inference, cut depth {cut}", flush=True)
    sc, bvh = small
    p = sq_amd.Proposal(bvh, cut_depth=cut, seed=0xC0 + cut)
    # non-zero biases so the last layer's bias path is exercised
    layers = p.weights()
    rng = np.random.default_rng(cut)
    layers = [(w, (0.1 * rng.standard_normal(b.shape)).astype(np.float32)) for (w, b) in layers]
    p.set_weights(layers)
    rays = W.proposal_training_rays(sc, 5000, seed=0x1FE + cut)
    top = p.infer(dev(rays)).cpu().numpy().view(np.uint16).astype(np.int64)
    n = p.n_subtrees
    B, box = p.features_spec()
    _, z = mlp.forward(mlp.features(rays, B, box), layers)
    z = np.asarray(z, np.float64)[:, :n]
    k = min(4, n)
    assert np.all(top[:, k:] == 0xFFFF)
    t = top[:, :k]
    assert np.all(t < n)
    assert all(len(set(row)) == k for row in t)
    best = -np.sort(-z, axis=1)[:, :k]
    picked = np.take_along_axis(z, t, 1)
    tol = 1e-3 * np.abs(z).max()       # fp32 accumulation order vs float64
    assert np.all(picked >= best - tol)
    assert (t == mlp.top4(z, n)[:, :k]).all(axis=1).mean() >= 0.98
    p.close()


@pytest.mark.parametrize("mode", ["untrained", "random_ids", "empty", "none", "short_tmax"])
def test_proposed_walk_is_exact(ctx, small, mode):
    import torch
    print(f"case: proposed walk, {mode}", flush=True)
    sc, bvh = small
    p = sq_amd.Proposal(bvh, cut_depth=9, seed=0x77)
    rays = np.concatenate([W.camera_rays(sc, res=256, seed=0x11), W.incoherent_rays(sc, 65536)])
    r = dev(rays)
    tmax = 40.0 if mode == "short_tmax" else 2000.0
    plain = bvh.raycast_nearest(r, tmax)
    if mode in ("untrained", "short_tmax"):
        top = p.infer(r)
    elif mode == "random_ids":
        g = torch.Generator(device="cuda"); g.manual_seed(5)
        top = torch.randint(0, p.n_subtrees, (r.shape[0], 4), device="cuda", generator=g).to(torch.int16)
    elif mode == "empty":
        top = torch.full((r.shape[0], 4), -1, dtype=torch.int16, device="cuda")
    else:
        top = None
    got = p.raycast_nearest(r, tmax, top, count=True)
    gl, gt = got["leaf"].cpu().numpy(), got["toi"].cpu().numpy()
    assert np.array_equal(gl, plain["leaf"].cpu().numpy())
    assert np.array_equal(gt.view(np.uint32), plain["toi"].cpu().numpy().view(np.uint32))
    if mode in ("random_ids", "short_tmax"):
        ref = oracle.Scene(3, oracle.TRI, sc.prims, guarded=True).raycast_nearest(rays, tmax, threads=8)
        assert_nearest_equal(gl, gt, ref["leaf"], ref["t"], f"proposed walk ({mode})")
    if mode in ("empty", "none"):
        assert got["seeded"] == 0
    # same work with the counters off
    fast = p.raycast_nearest(r, tmax, top)
    assert np.array_equal(fast["leaf"].cpu().numpy(), gl)
    p.close()


def test_training_gains_over_untrained(ctx, small):
    """Train on fresh draws from the evaluation distributions (jittered
    camera rays of the same camera, incoherent rays of another seed); on the
    evaluation camera rays (pixel centres) the trained network proposes the hit
    subtree far more often than the untrained one, the proposals seed many
    more rays, and the results stay exact.  Incoherent rays are reported, not
    asserted: their hit subtree is not a function the network learns from
    this much data."""
    print("case: proposal training gain", flush=True)
    sc, bvh = small
    p = sq_amd.Proposal(bvh, cut_depth=8, rff_scale=1.0, seed=0x3)
    train = W.proposal_training_rays(sc, 1 << 18, seed=0x7A1)
    rt, lab, _ = _labelled(bvh, p, train)
    n_cam = 256 * 256
    evalr = np.concatenate([W.camera_rays(sc, res=256), W.incoherent_rays(sc, 65536, seed=0xE7A1)])
    re, lab_e, exact = _labelled(bvh, p, evalr)
    lab_e = lab_e.cpu().numpy()

    def hit_rate(top, sl):
        t = top.cpu().numpy().view(np.uint16).astype(np.int64)[sl]
        l = lab_e[sl]
        ok = l >= 0
        return float((t[ok] == l[ok, None]).any(axis=1).mean())

    cam, inc = slice(0, n_cam), slice(n_cam, None)
    assert (lab_e[cam] >= 0).mean() > 0.4
    top0 = p.infer(re)
    h0 = hit_rate(top0, cam)
    w0 = p.raycast_nearest(re, 2000.0, top0, count=True)
    losses = p.fit(rt, lab, epochs=8, batch=8192, lr=1e-3)
    top1 = p.infer(re)
    h1, h1i = hit_rate(top1, cam), hit_rate(top1, inc)
    w1 = p.raycast_nearest(re, 2000.0, top1, count=True)
    plain = p.raycast_nearest(re, 2000.0, None, count=True)
    n = re.shape[0]
    print(f"subtrees {p.n_subtrees}; loss {losses[0]:.3f} -> {losses[-1]:.3f}; camera top-4 hit rate {h0:.3f} -> {h1:.3f} "
          f"(incoherent {h1i:.3f}); seeded rays {w0['seeded']} -> {w1['seeded']}; "
          f"nodes/ray plain {plain['nodes'] / n:.2f}, untrained {w0['nodes'] / n:.2f}, trained {w1['nodes'] / n:.2f}; "
          f"prims/ray plain {plain['prims'] / n:.2f}, untrained {w0['prims'] / n:.2f}, trained {w1['prims'] / n:.2f}",
          flush=True)
    assert losses[-1] < losses[0]
    assert h1 >= 0.8 and h1 >= h0 + 0.5
    assert w1["seeded"] > 2 * w0["seeded"]
    assert np.array_equal(w1["leaf"].cpu().numpy(), exact["leaf"].cpu().numpy())
    assert np.array_equal(w1["toi"].cpu().numpy().view(np.uint32), exact["toi"].cpu().numpy().view(np.uint32))
    p.close()
