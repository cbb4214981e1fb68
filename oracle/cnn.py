"""CPU restatement of the learned centroid proposal -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` and ``bench.py``'s checking legs import this module, to check
the HIP kernels against it; the product never imports it.

The reference has no learned path: this restates the network the learned
proposal is specified as, in numpy.

* Input: the 11 x 11 patch of the bandpassed frame centred on a candidate
  pixel.  Pixels outside the frame are 0.  The patch is divided by its
  largest absolute value (a 0 patch stays 0), with a float32 division, and
  rounded to bfloat16.
* conv3x3 1->16, conv3x3 16->32, conv3x3 32->32.  These are valid
  cross-correlations (11 -> 9 -> 7 -> 5), each with a bias and a ReLU.
  Weights are rounded to bfloat16 and biases stay float32.  Each activation
  is accumulated in float32 and rounded to bfloat16.
* Global average pooling over the 5 x 5 positions (sum / 25, rounded to
  bfloat16), then a linear layer 32 -> 3 with float32 output:
  (dy, dx, confidence logit).  dy is along rows and dx along columns,
  relative to the candidate pixel.
* Loss per sample: label * ((dy - ty)^2 + (dx - tx)^2) + BCE(logit, label),
  averaged over the batch; the position term is 0 where label is 0 (any
  target offset, even a non-finite one).  (ty, tx) is the walked centroid
  minus the candidate pixel; label is 1 when the walk converges.
* Backward, with the same rounding points as the kernels.  The gradients at
  the pre-activations of conv3 and conv2 are rounded to bfloat16, the inputs
  of the matrix-core products.  The conv1 gradient stays float32.  The ReLU
  derivative is (activation > 0).
* Adam in torch's form (bias-corrected; eps added to sqrt(v) / sqrt(bc2)).

Parameter vector (float32, 14147 entries, torch layouts, in this order):
w1 (16,1,3,3), b1 (16), w2 (32,16,3,3), b2 (32), w3 (32,32,3,3), b3 (32),
wfc (3,32), bfc (3).

Pinning (tests/test_oracle_cnn.py):

* With rounding off, forward and gradients are checked against torch
  autograd in float64 and against central differences.
* bf16() is checked against torch's bfloat16 conversion.
* adam_step() is checked against torch.optim.Adam.
"""
from __future__ import annotations

import numpy as np

PATCH = 11
NPARAM = 14147
LAYOUT = [
    ("w1", (16, 1, 3, 3)), ("b1", (16,)), ("w2", (32, 16, 3, 3)), ("b2", (32,)),
    ("w3", (32, 32, 3, 3)), ("b3", (32,)), ("wfc", (3, 32)), ("bfc", (3,)),
]


def offsets():
    out, o = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = (o, shape)
        o += n
    assert o == NPARAM
    return out


OFF = offsets()


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even); float32 out."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(params):
    p = np.asarray(params, np.float32)
    assert p.shape == (NPARAM,)
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OFF.items()}


def pack(d):
    p = np.zeros(NPARAM, np.float32)
    for k, (o, s) in OFF.items():
        p[o:o + int(np.prod(s))] = np.asarray(d[k], np.float32).reshape(-1)
    return p


def gather_patches(proc, cands):
    """Raw 11 x 11 patches (n, 121) float32 of the processed frame `proc`
    (H, W) around candidate pixels `cands` (n, 2) = (row, col); 0 outside."""
    proc = np.asarray(proc, np.float32)
    H, W = proc.shape
    pad = np.zeros((H + 10, W + 10), np.float32)
    pad[5:5 + H, 5:5 + W] = proc
    c = np.asarray(cands, np.int64).reshape(-1, 2)
    out = np.empty((len(c), PATCH * PATCH), np.float32)
    for i, (u, v) in enumerate(c):
        out[i] = pad[u:u + PATCH, v:v + PATCH].reshape(-1)
    return out


def normalize_patches(raw):
    """(n, 121) raw patches -> (n, 11, 11) network inputs (bfloat16 values)."""
    raw = np.asarray(raw, np.float32).reshape(-1, PATCH * PATCH)
    s = np.abs(raw).max(axis=1, keepdims=True)
    with np.errstate(divide="ignore", invalid="ignore"):
        q = np.where(s > 0, raw / np.where(s > 0, s, np.float32(1)), np.float32(0))
    return bf16(q.astype(np.float32)).reshape(-1, PATCH, PATCH)


def _conv(x, w, b):
    """Valid 3x3 cross-correlation, (n, ci, h, w) -> (n, co, h-2, w-2), float64."""
    n, ci, h, wd = x.shape
    out = np.zeros((n, w.shape[0], h - 2, wd - 2), np.float64)
    for ky in range(3):
        for kx in range(3):
            out += np.einsum("ncij,oc->noij", x[:, :, ky:ky + h - 2, kx:kx + wd - 2], w[:, :, ky, kx])
    return out + b[None, :, None, None]


def _round(x, rounding):
    x = np.asarray(x, np.float64)
    if not rounding:
        return x
    return bf16(x.astype(np.float32)).astype(np.float64)


def forward(params, x, rounding=True, keep=False):
    """x: (n, 11, 11) network inputs. Returns (n, 3) float64 [dy, dx, logit]
    (and the activations when keep)."""
    P = {k: np.asarray(v, np.float64) for k, v in unpack(params).items()}
    for k in ("w1", "w2", "w3", "wfc"):
        P[k] = _round(P[k], rounding)
    x = np.asarray(x, np.float64).reshape(-1, 1, PATCH, PATCH)
    a1 = _round(np.maximum(_conv(x, P["w1"], P["b1"]), 0), rounding)
    a2 = _round(np.maximum(_conv(a1, P["w2"], P["b2"]), 0), rounding)
    a3 = _round(np.maximum(_conv(a2, P["w3"], P["b3"]), 0), rounding)
    g = _round(a3.sum(axis=(2, 3)) / 25.0, rounding)
    o = g @ P["wfc"].T + P["bfc"][None, :]
    if keep:
        return o, dict(P=P, x=x, a1=a1, a2=a2, a3=a3, g=g, o=o)
    return o


def loss(o, targets):
    """Batch loss of outputs o (n, 3) against targets (n, >=3): ty, tx, label."""
    t = np.asarray(targets, np.float64)
    ty, tx, lab = t[:, 0], t[:, 1], t[:, 2]
    on = lab != 0
    reg = np.where(on, lab * ((o[:, 0] - ty) ** 2 + (o[:, 1] - tx) ** 2), 0.0)
    z = o[:, 2]
    bce = np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z))) - lab * z
    return float(np.mean(reg + bce))


def gradients(params, x, targets, rounding=True):
    """Gradient of loss(forward(params, x), targets) with respect to the
    parameter vector, float64 (NPARAM,), and the loss."""
    o, A = forward(params, x, rounding, keep=True)
    P = A["P"]
    n = o.shape[0]
    t = np.asarray(targets, np.float64)
    ty, tx, lab = t[:, 0], t[:, 1], t[:, 2]
    on = lab != 0
    do = np.zeros_like(o)
    do[:, 0] = np.where(on, lab * 2.0 * (o[:, 0] - ty), 0.0) / n
    do[:, 1] = np.where(on, lab * 2.0 * (o[:, 1] - tx), 0.0) / n
    do[:, 2] = (1.0 / (1.0 + np.exp(-o[:, 2])) - lab) / n
    G = {}
    G["wfc"] = do.T @ A["g"]
    G["bfc"] = do.sum(0)
    dg = do @ P["wfc"]                                      # (n, 32)
    da3 = np.broadcast_to((dg / 25.0)[:, :, None, None], A["a3"].shape)
    dz3 = _round(da3 * (A["a3"] > 0), rounding)
    G["b3"] = dz3.sum(axis=(0, 2, 3))
    G["w3"] = np.zeros((32, 32, 3, 3))
    da2 = np.zeros(A["a2"].shape)
    for ky in range(3):
        for kx in range(3):
            G["w3"][:, :, ky, kx] = np.einsum("noij,ncij->oc", dz3, A["a2"][:, :, ky:ky + 5, kx:kx + 5])
            da2[:, :, ky:ky + 5, kx:kx + 5] += np.einsum("noij,oc->ncij", dz3, P["w3"][:, :, ky, kx])
    dz2 = _round(da2 * (A["a2"] > 0), rounding)
    G["b2"] = dz2.sum(axis=(0, 2, 3))
    G["w2"] = np.zeros((32, 16, 3, 3))
    da1 = np.zeros(A["a1"].shape)
    for ky in range(3):
        for kx in range(3):
            G["w2"][:, :, ky, kx] = np.einsum("noij,ncij->oc", dz2, A["a1"][:, :, ky:ky + 7, kx:kx + 7])
            da1[:, :, ky:ky + 7, kx:kx + 7] += np.einsum("noij,oc->ncij", dz2, P["w2"][:, :, ky, kx])
    dz1 = da1 * (A["a1"] > 0)
    G["b1"] = dz1.sum(axis=(0, 2, 3))
    G["w1"] = np.zeros((16, 1, 3, 3))
    for ky in range(3):
        for kx in range(3):
            G["w1"][:, :, ky, kx] = np.einsum("noij,ncij->oc", dz1, A["x"][:, :, ky:ky + 9, kx:kx + 9])
    flat = np.zeros(NPARAM, np.float64)
    for k, (off, s) in OFF.items():
        flat[off:off + int(np.prod(s))] = np.asarray(G[k]).reshape(-1)
    return flat, loss(o, targets)


def adam_step(p, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (torch.optim.Adam's arithmetic, float32). Returns p, m, v."""
    p, g, m, v = (np.asarray(a, np.float32) for a in (p, g, m, v))
    m = (np.float32(b1) * m + np.float32(1 - b1) * g).astype(np.float32)
    v = (np.float32(b2) * v + np.float32(1 - b2) * g * g).astype(np.float32)
    bc1 = np.float32(1 - b1 ** t)
    bc2 = np.float32(1 - b2 ** t)
    denom = np.sqrt(v) / np.sqrt(bc2) + np.float32(eps)
    p = (p - np.float32(lr) / bc1 * m / denom).astype(np.float32)
    return p, m, v


def init_params(seed):
    """He-uniform weights (bound sqrt(6 / fan_in)), zero biases; the same
    generator as the product's initialiser is not required: tests pass
    parameters explicitly."""
    rng = np.random.default_rng(seed)
    d = {}
    for k, s in LAYOUT:
        if k.startswith("b"):
            d[k] = np.zeros(s, np.float32)
        else:
            fan_in = int(np.prod(s[1:]))
            d[k] = rng.uniform(-1, 1, s).astype(np.float32) * np.float32(np.sqrt(6.0 / fan_in))
    return pack(d)
