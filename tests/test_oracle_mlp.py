"""Pins the MLP oracle (oracle/mlp.py) that checks the device proposal network.

The network has no counterpart in the reference, so it is pinned by what the
mathematics fixes: bfloat16 rounding against torch's own conversion and the
ties-to-even special cases, the features' sin^2 + cos^2 = 1 and their value at
the origin, the loss against scipy's log_softmax, the backward pass against
central finite differences in float64 (a dropped term, a wrong sign or a
transposed operand fails these), and top-4 against a brute-force sort.
"""
import numpy as np
import pytest
import scipy.special
import torch

from oracle import mlp


def test_bf16_round_special_cases():
    one = np.float32(1.0)
    cases = {
        1.0: 1.0,
        1.0 + 2.0 ** -8: 1.0,                     # tie -> even (mantissa 0)
        1.0 + 3 * 2.0 ** -8: 1.0 + 2.0 ** -6,     # tie -> even (mantissa 2)
        1.0 + 2.0 ** -8 + 2.0 ** -20: 1.0 + 2.0 ** -7,
        -2.5: -2.5,
        0.0: 0.0,
    }
    for x, want in cases.items():
        assert mlp.bf16_round(np.float32(x)) == np.float32(want), x
    assert np.isinf(mlp.bf16_round(np.float32(np.inf)))
    assert np.isnan(mlp.bf16_round(np.float32(np.nan)))
    assert one == mlp.bf16_round(one)


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.standard_normal(100_000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 100_000),
                        rng.standard_normal(1000).astype(np.float32) * 1e-39]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(mlp.bf16_round(x).view(np.uint32), want.view(np.uint32))


def test_features_identities():
    rng = np.random.default_rng(3)
    B = (3.0 * rng.standard_normal((6, 64))).astype(np.float32)
    box = np.array([-10, -20, -5, 10, 20, 15], np.float32)
    rays = np.concatenate([rng.uniform(-10, 10, (500, 3)), rng.standard_normal((500, 3))], axis=1).astype(np.float32)
    F = mlp.features(rays, B, box, bf16=False)
    assert F.shape == (500, 128)
    assert np.allclose(F[:, :64] ** 2 + F[:, 64:] ** 2, 1.0, atol=1e-12)
    # box centre, zero direction: every phase is 0
    c = np.array([[0, 0, 5, 0, 0, 0]], np.float32)
    F0 = mlp.features(c, B, box, bf16=False)
    assert np.all(F0[0, :64] == 0.0) and np.all(F0[0, 64:] == 1.0)
    # float64 restatement of the phase within float32 rounding
    x = np.concatenate([(rays[:, :3].astype(np.float64) - box[:3]) / (box[3:] - box[:3]) * 2 - 1, rays[:, 3:]], axis=1)
    a = 2 * np.pi * (x @ B.astype(np.float64))
    assert np.allclose(F[:, :64], np.sin(a), atol=1e-4)
    Fb = mlp.features(rays, B, box, bf16=True)
    assert np.array_equal(Fb, mlp.bf16_round(F.astype(np.float32)))


def _small_net(rng, widths):
    layers = []
    for i in range(4):
        W = rng.standard_normal((widths[i + 1], widths[i])) * np.sqrt(2.0 / widths[i])
        b = rng.standard_normal(widths[i + 1]) * 0.1
        layers.append((W, b))
    return layers


def test_loss_matches_log_softmax():
    rng = np.random.default_rng(11)
    z = rng.standard_normal((40, 9)) * 3
    lab = rng.integers(-1, 9, 40)
    loss, dz = mlp.softmax_ce(z, lab, bf16=False)
    ls = scipy.special.log_softmax(z, axis=1)
    ok = lab >= 0
    assert loss == pytest.approx(-ls[ok, lab[ok]].mean(), rel=1e-12)
    assert np.all(dz[~ok] == 0.0)
    assert np.allclose(dz[ok].sum(axis=1), 0.0, atol=1e-15)


@pytest.mark.parametrize("seed", [1, 2])
def test_backward_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    widths = [5, 7, 6, 8, 9]
    layers = _small_net(rng, widths)
    F = rng.standard_normal((12, 5))
    lab = rng.integers(0, 9, 12)
    lab[3] = -1
    loss, grads = mlp.loss_and_grads(F, layers, lab, bf16=False)

    def f(ls):
        return mlp.loss_and_grads(F, ls, lab, bf16=False)[0]

    h = 1e-6
    for l in range(4):
        for which in (0, 1):
            P = layers[l][which]
            G = grads[l][which]
            flat = P.reshape(-1)
            for idx in rng.choice(flat.size, size=min(flat.size, 12), replace=False):
                old = flat[idx]
                flat[idx] = old + h; lp = f(layers)
                flat[idx] = old - h; lm = f(layers)
                flat[idx] = old
                fd = (lp - lm) / (2 * h)
                assert G.reshape(-1)[idx] == pytest.approx(fd, rel=1e-5, abs=1e-9), (l, which, idx)


def test_bf16_mode_close_to_float64():
    rng = np.random.default_rng(5)
    widths = [128, 256, 256, 256, 40]
    layers = _small_net(rng, widths)
    F = mlp.bf16_round(rng.uniform(-1, 1, (256, 128)).astype(np.float32))
    lab = rng.integers(0, 40, 256)
    _, g64 = mlp.loss_and_grads(F, layers, lab, bf16=False)
    _, g16 = mlp.loss_and_grads(F, layers, lab, bf16=True)
    # bf16 weights and activations perturb the gradients by a few percent,
    # growing with the number of layers the error is propagated through
    # (measured: 0.4% for the last layer, 6.6% for the first)
    for (a, b), (c, d) in zip(g64, g16):
        assert np.linalg.norm(a - c) <= 0.1 * np.linalg.norm(a)
        assert np.linalg.norm(b - d) <= 0.1 * np.linalg.norm(b)


def test_top4_brute_force():
    rng = np.random.default_rng(9)
    z = rng.integers(-3, 4, (200, 30)).astype(np.float32)      # many ties
    got = mlp.top4(z, 25)
    for r in range(200):
        order = sorted(range(25), key=lambda c: (-z[r, c], c))[:4]
        assert list(got[r]) == order
