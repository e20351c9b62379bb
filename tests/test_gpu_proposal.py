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


@pytest.mark.parametrize("group", [3, 5])
def test_infer_top4_exact_ties_keep_the_smaller_ids(ctx, small, group):
    """Exact ties in the fused top-4: the last layer's rows and biases repeat
    in groups of 3 or 5 classes (so a group's classes straddle the lanes and
    partial lists that hold them), making the logits of a group bit-identical
    on the device.  The top four are then the classes of the best groups in
    ascending id order: for every ray whose two best groups are separated
    beyond the float64 oracle's rounding tolerance, the device's list is
    exactly that (a tie kept in the wrong order, or a smaller id lost to a
    later tie, fails)."""
    print(f"case: fused top-4 with exact ties, groups of {group}", flush=True)
    sc, bvh = small
    p = sq_amd.Proposal(bvh, cut_depth=9, seed=0x7E + group)
    layers = p.weights()
    rng = np.random.default_rng(group)
    n = p.n_subtrees
    w4, b4 = layers[-1]
    g = np.arange(w4.shape[0]) // group
    w4 = w4[g * group].copy()
    b4 = (0.2 * rng.standard_normal(b4.shape)).astype(np.float32)[g * group]
    layers = layers[:-1] + [(w4, b4)]
    p.set_weights(layers)
    rays = W.proposal_training_rays(sc, 20000, seed=0x71E + group)
    top = p.infer(dev(rays)).cpu().numpy().view(np.uint16).astype(np.int64)
    B, box = p.features_spec()
    _, z = mlp.forward(mlp.features(rays, B, box), layers)
    z = np.asarray(z, np.float64)[:, :n]
    zg = z[:, ::group]                              # one logit per group
    order = np.argsort(-zg, axis=1, kind="stable")
    tol = 1e-3 * np.abs(z).max()
    sizes = np.minimum(group, n - np.arange(zg.shape[1]) * group)
    checked = 0
    for i in range(len(rays)):
        want, k = [], 0
        while len(want) < 4:
            gi = order[i, k]
            want += list(range(gi * group, gi * group + sizes[gi]))[: 4 - len(want)]
            k += 1
        # the groups that supply the four must be separated from the next ones
        if zg[i, order[i, k - 1]] - zg[i, order[i, k]] <= tol or any(
                zg[i, order[i, j]] - zg[i, order[i, j + 1]] <= tol for j in range(k - 1)):
            continue
        assert top[i].tolist() == want, (i, top[i].tolist(), want)
        checked += 1
    assert checked > 10000
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


# --- the proposal over a two-level scene (Tlas top tree cut into subtrees) ---

@pytest.fixture(scope="module")
def instanced(ctx):
    sc = W.instanced_scene(n_meshes=5, n_inst=400, tri_lo=200, tri_hi=3000, extent=100.0)
    meshes = [sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, m) for m in sc["meshes"]]
    t = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.MESH, sc["rot"], sc["trans"], meshes=meshes, mesh_of=sc["mesh_of"], scale=sc["scale"])
    yield sc, t
    t.close()
    for m in meshes:
        m.close()


def _top_for(p, r, mode):
    import torch
    if mode in ("untrained", "short_tmax"):
        return p.infer(r)
    if mode == "random_ids":
        g = torch.Generator(device="cuda"); g.manual_seed(9)
        return torch.randint(0, p.n_subtrees, (r.shape[0], 4), device="cuda", generator=g).to(torch.int16)
    if mode == "empty":
        return torch.full((r.shape[0], 4), -1, dtype=torch.int16, device="cuda")
    return None


@pytest.mark.parametrize("mode", ["untrained", "random_ids", "empty", "none", "short_tmax"])
def test_tlas_proposed_walk_is_exact(ctx, instanced, mode):
    """Mesh instances: the seeded two-level walk returns the guarded
    sq_tlas_raycast_nearest's instance, toi and triangle bit for bit whatever
    the proposals, and the guarded CPU oracle's."""
    print(f"case: tlas proposed walk, {mode}", flush=True)
    sc, t = instanced
    p = sq_amd.Proposal(t, cut_depth=6, seed=0x78)
    assert 16 <= p.n_subtrees <= 64
    rays = W.random_rays(30000, 0xC3, -120.0, 120.0)
    r = dev(rays)
    tmax = 40.0 if mode == "short_tmax" else 1000.0
    plain = t.raycast_nearest(r, tmax, sub=True)
    got = p.raycast_nearest(r, tmax, _top_for(p, r, mode), count=True)
    gl, gt, gs = (got[k].cpu().numpy() for k in ("leaf", "toi", "sub"))
    assert np.array_equal(gl, plain["leaf"].cpu().numpy())
    assert np.array_equal(gt.view(np.uint32), plain["toi"].cpu().numpy().view(np.uint32))
    assert np.array_equal(gs, plain["sub"].cpu().numpy())
    hit = gl >= 0
    assert hit.sum() > 200
    if mode in ("random_ids", "short_tmax"):
        o = oracle.InstanceScene(3, oracle.INST_MESH, sc["rot"], sc["trans"], mesh_tris=sc["meshes"], mesh_of=sc["mesh_of"],
                                 scale=sc["scale"], guarded=True)
        ref = o.raycast_nearest(rays, tmax, threads=8)
        assert_nearest_equal(gl, gt, ref["leaf"], ref["t"], f"tlas proposed walk ({mode})")
        assert np.array_equal(gs[hit], ref["sub"][hit])
    if mode in ("empty", "none"):
        assert got["seeded"] == 0
    if mode == "random_ids":
        assert got["seeded"] > 0
    p.close()


@pytest.mark.parametrize("kind", ["box", "ball"])
def test_tlas_proposed_walk_box_and_ball_instances(ctx, kind):
    print(f"case: tlas proposed walk, {kind} instances", flush=True)
    n = 3000
    s = W.Stream(0xB0B + (kind == "ball"))
    rot = W.random_rotations(s, n)
    tr = s.range(-100.0, 100.0, n * 3).reshape(n, 3)
    if kind == "box":
        halves = s.range(0.3, 3.0, n * 3).reshape(n, 3)
        t = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.BOX, rot, tr, halves=halves)
        o = oracle.InstanceScene(3, oracle.INST_BOX, rot, tr, halves=halves, guarded=True)
    else:
        halves = s.range(0.3, 3.0, n)
        t = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.BALL, rot, tr, halves=halves)
        o = oracle.InstanceScene(3, oracle.INST_BALL, rot, tr, halves=halves, guarded=True)
    p = sq_amd.Proposal(t, cut_depth=9, seed=0x79)
    rays = W.random_rays(40000, 0xC4, -120.0, 120.0)
    r = dev(rays)
    plain = t.raycast_nearest(r, 1000.0)
    got = p.raycast_nearest(r, 1000.0, _top_for(p, r, "random_ids"))
    gl, gt = got["leaf"].cpu().numpy(), got["toi"].cpu().numpy()
    assert np.array_equal(gl, plain["leaf"].cpu().numpy())
    assert np.array_equal(gt.view(np.uint32), plain["toi"].cpu().numpy().view(np.uint32))
    ref = o.raycast_nearest(rays, 1000.0, threads=8)
    assert (ref["leaf"] >= 0).sum() > 2000
    assert_nearest_equal(gl, gt, ref["leaf"], ref["t"], f"tlas proposed walk ({kind} instances)")
    p.close()
    t.close()


def _cone_rays(n, seed, origin, half_angle_deg):
    """Rays from one origin, directions uniform in the cone of the given half
    angle around the direction to the scene's centre (the world origin)."""
    s = W.Stream(seed)
    u = s.range(0.0, 1.0, 2 * n).reshape(n, 2).astype(np.float64)
    o = np.asarray(origin, np.float64)
    f = -o / np.linalg.norm(o)
    a = np.array([1.0, 0.0, 0.0]) if abs(f[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(f, a); e1 /= np.linalg.norm(e1)
    e2 = np.cross(f, e1)
    ct = 1.0 - u[:, 0] * (1.0 - np.cos(np.radians(half_angle_deg)))
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    ph = 2.0 * np.pi * u[:, 1]
    d = ct[:, None] * f + (st * np.cos(ph))[:, None] * e1 + (st * np.sin(ph))[:, None] * e2
    d = W.normalize_or_zero(d.astype(np.float32))
    return W.rays_array(np.broadcast_to(o.astype(np.float32), d.shape), d)


def test_tlas_training_gains_over_untrained(ctx, instanced):
    """Rays from one viewpoint into the instance cloud (training and
    evaluation directions drawn from the same cone with different seeds);
    labels are the top-tree subtrees of the hit instances.  The trained
    network proposes the hit subtree far more often than the untrained one,
    seeds more rays, and the results stay exact."""
    print("case: tlas proposal training gain", flush=True)
    sc, t = instanced
    p = sq_amd.Proposal(t, cut_depth=6, rff_scale=1.0, seed=0x4)
    eye = (60.0, -230.0, 90.0)
    train = _cone_rays(1 << 18, 0x7A2, eye, 25.0)
    rt = dev(train)
    lab = p.labels(t.raycast_nearest(rt, 1000.0)["leaf"])
    evalr = _cone_rays(1 << 16, 0xE7A2, eye, 25.0)
    re = dev(evalr)
    exact = t.raycast_nearest(re, 1000.0, sub=True)
    lab_e = p.labels(exact["leaf"]).cpu().numpy()
    ok = lab_e >= 0

    def hit_rate(top):
        tt = top.cpu().numpy().view(np.uint16).astype(np.int64)
        return float((tt[ok] == lab_e[ok, None]).any(axis=1).mean())

    top0 = p.infer(re)
    h0 = hit_rate(top0)
    w0 = p.raycast_nearest(re, 1000.0, top0, count=True)
    losses = p.fit(rt, lab, epochs=8, batch=8192, lr=1e-3)
    top1 = p.infer(re)
    h1 = hit_rate(top1)
    w1 = p.raycast_nearest(re, 1000.0, top1, count=True)
    print(f"subtrees {p.n_subtrees}; labelled eval rays {int(ok.sum())}; loss {losses[0]:.3f} -> {losses[-1]:.3f}; "
          f"top-4 hit rate {h0:.3f} -> {h1:.3f}; seeded rays {w0['seeded']} -> {w1['seeded']}", flush=True)
    assert ok.sum() > 5000
    assert losses[-1] < losses[0]
    assert h1 >= h0 + 0.3
    assert w1["seeded"] > 2 * w0["seeded"]
    for k in ("leaf", "sub"):
        assert np.array_equal(w1[k].cpu().numpy(), exact[k].cpu().numpy())
    assert np.array_equal(w1["toi"].cpu().numpy().view(np.uint32), exact["toi"].cpu().numpy().view(np.uint32))
    p.close()


def test_c3_configuration_heightfield_cut12(ctx):
    """The learned proposal at its own configuration on the full 2,097,152-
    triangle heightfield: cut depth 12 (<= 4,096 subtrees), a short Adam run
    (1M labelled training rays, lr 1e-3, batch 65,536), then 1M fresh
    evaluation rays from the heightfield distributions (half camera rays at
    pixel centres, half incoherent): the seeded walk equals the plain walk on
    every ray and the guarded CPU oracle on a sample, and training raises the
    camera rays' top-4 hit rate."""
    print("case: C3 configuration, 2M-triangle heightfield, cut depth 12", flush=True)
    import torch
    sc = W.heightfield_scene()
    bvh = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, sc.prims)
    p = sq_amd.Proposal(bvh, cut_depth=12, rff_scale=1.0, seed=0xC3)
    assert 2048 < p.n_subtrees <= 4096
    train = W.proposal_training_rays(sc, 1 << 20, seed=0xC3C3)
    rt, lab, _ = _labelled(bvh, p, train)
    cam = W.camera_rays(sc, res=2048)[::8]                       # 524,288 pixel-centre rays
    inc = W.incoherent_rays(sc, 1 << 19, seed=0xC3E0)
    evalr = np.ascontiguousarray(np.concatenate([cam, inc]))
    re, lab_e, exact = _labelled(bvh, p, evalr)
    lab_e = lab_e.cpu().numpy()
    sl = slice(0, cam.shape[0])

    def hit_rate(top):
        tt = top.cpu().numpy().view(np.uint16).astype(np.int64)[sl]
        l = lab_e[sl]
        ok = l >= 0
        return float((tt[ok] == l[ok, None]).any(axis=1).mean())

    h0 = hit_rate(p.infer(re))
    losses = p.fit(rt, lab, epochs=1, batch=65536, lr=1e-3)
    top = p.infer(re)
    h1 = hit_rate(top)
    got = p.raycast_nearest(re, 2000.0, top, count=True)
    torch.cuda.synchronize()
    n = re.shape[0]
    print(f"subtrees {p.n_subtrees}; loss {losses[0]:.3f}; camera top-4 hit rate {h0:.3f} -> {h1:.3f}; "
          f"seeded {got['seeded']} of {n}", flush=True)
    gl, gt = got["leaf"].cpu().numpy(), got["toi"].cpu().numpy()
    assert np.array_equal(gl, exact["leaf"].cpu().numpy())
    assert np.array_equal(gt.view(np.uint32), exact["toi"].cpu().numpy().view(np.uint32))
    idx = np.concatenate([np.arange(0, cam.shape[0], 256), cam.shape[0] + np.arange(0, inc.shape[0], 256)])
    ref = oracle.Scene(3, oracle.TRI, sc.prims, guarded=True).raycast_nearest(evalr[idx], 2000.0, threads=8)
    assert_nearest_equal(gl[idx], gt[idx], ref["leaf"], ref["t"], "C3 proposed walk vs oracle sample")
    assert (ref["leaf"] >= 0).mean() > 0.5
    assert h1 > h0 + 0.2
    p.close()
    bvh.close()
