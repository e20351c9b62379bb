"""CPU restatement of the learned subtree proposal's network -- TEST INFRASTRUCTURE ONLY.

The proposal has no counterpart in the reference (its walk is the plain one,
crates/spatial-query/src/accel/bvh.rs:51-124); this module is the checker for
the device MLP: forward pass, softmax cross-entropy, backward pass and top-4
inference, in numpy.  Only tests may import it.

Two modes:
  * bf16=True  -- rounds to bfloat16 (round to nearest even) at exactly the
    points the device path stores bf16: features, hidden activations, the
    logit gradient dZ and the hidden gradients dH; weights are used as their
    bf16 roundings; products accumulate in float64 and are rounded to float32
    (the device accumulates in float32 in another order, so agreement is to a
    tolerance, stated in the tests);
  * bf16=False -- plain float64 everywhere: the mode the finite-difference
    tests pin the backward pass in.

Layers: h1 = relu(F W1^T + b1), h2, h3 likewise, z = h3 W4^T + b4;
loss = mean over labelled rows of -log softmax(z)[label].
"""
from __future__ import annotations

import numpy as np

N_RFF = 64
TWO_PI_F32 = np.float32(6.28318530717958647692)


def bf16_round(x) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = np.where(nan, a, out)
    return out.reshape(a.shape)


def ray_inputs(rays: np.ndarray, box: np.ndarray) -> np.ndarray:
    """x = ((o - lo) / (hi - lo) * 2 - 1, d) in float32, operation by operation."""
    r = np.asarray(rays, np.float32)
    lo = np.asarray(box[:3], np.float32)
    hi = np.asarray(box[3:], np.float32)
    o = ((r[:, :3] - lo) / (hi - lo)) * np.float32(2.0) - np.float32(1.0)
    return np.concatenate([o, r[:, 3:6]], axis=1).astype(np.float32)


def features(rays: np.ndarray, B: np.ndarray, box: np.ndarray, bf16: bool = True) -> np.ndarray:
    """Random Fourier features [n][128]: sin(2 pi x.B) then cos(2 pi x.B).

    x.B accumulates k = 0..5 in float32 left to right, as the device does."""
    x = ray_inputs(rays, box)
    Bf = np.asarray(B, np.float32)
    z = np.zeros((x.shape[0], N_RFF), np.float32)
    for k in range(6):
        z = (z + x[:, k:k + 1] * Bf[k][None, :]).astype(np.float32)
    a = (TWO_PI_F32 * z).astype(np.float32)
    F = np.concatenate([np.sin(a.astype(np.float64)), np.cos(a.astype(np.float64))], axis=1)
    return bf16_round(F.astype(np.float32)) if bf16 else F


def _mm(a, b, bf16):
    """a @ b accumulated in float64; rounded to float32 in bf16 mode."""
    c = np.asarray(a, np.float64) @ np.asarray(b, np.float64)
    return c.astype(np.float32) if bf16 else c


def forward(F, layers, bf16: bool = True):
    """Hidden activations [h1, h2, h3] and logits z for features F; layers = [(W, b)] x 4."""
    R = bf16_round if bf16 else (lambda v: np.asarray(v, np.float64))
    Ws = [R(W) for (W, _) in layers]
    bs = [np.asarray(b, np.float32 if bf16 else np.float64) for (_, b) in layers]
    hs = []
    x = F
    for l in range(3):
        pre = _mm(x, Ws[l].T, bf16) + bs[l]
        x = R(np.maximum(pre, 0.0))
        hs.append(x)
    z = _mm(x, Ws[3].T, bf16) + bs[3]
    return hs, z


def softmax_ce(z, labels, bf16: bool = True):
    """(mean loss, dZ) over labelled rows (label >= 0); dZ = (p - onehot) / count."""
    z = np.asarray(z, np.float64)
    lab = np.asarray(labels)
    m = z.max(axis=1, keepdims=True)
    e = np.exp(z - m)
    se = e.sum(axis=1, keepdims=True)
    p = e / se
    ok = lab >= 0
    cnt = int(ok.sum())
    inv = 1.0 / cnt if cnt else 0.0
    onehot = np.zeros_like(p)
    onehot[np.nonzero(ok)[0], lab[ok]] = 1.0
    dz = (p - onehot) * inv
    dz[~ok] = 0.0
    rows = np.nonzero(ok)[0]
    logp = (z - m - np.log(se))[rows, lab[ok]]
    loss = float(-logp.sum() * inv) if cnt else 0.0
    if bf16:
        dz = bf16_round(dz.astype(np.float32))
    return loss, dz


def loss_and_grads(F, layers, labels, bf16: bool = True):
    """Mean cross-entropy and its gradients [(dW, db)] x 4, with the device's rounding points."""
    R = bf16_round if bf16 else (lambda v: np.asarray(v, np.float64))
    Ws = [R(W) for (W, _) in layers]
    hs, z = forward(F, layers, bf16)
    loss, dz = softmax_ce(z, labels, bf16)
    xs = [F] + hs            # inputs of layers 1..4
    grads = [None] * 4
    d = dz
    for l in range(3, -1, -1):
        dW = _mm(np.asarray(d).T, xs[l], bf16)
        db = np.asarray(d, np.float64).sum(axis=0)
        grads[l] = (dW, db.astype(np.float32) if bf16 else db)
        if l > 0:
            dh = _mm(d, Ws[l], bf16)
            dh = np.where(np.asarray(xs[l]) > 0, dh, 0.0)
            d = R(dh.astype(np.float32)) if bf16 else dh
    return loss, grads


def top4(z, n_cls: int):
    """Best four classes per row: larger logit first, then the smaller class id."""
    z = np.asarray(z)[:, :n_cls]
    order = np.lexsort((np.broadcast_to(np.arange(n_cls), z.shape), -z), axis=1)
    return order[:, :4]
