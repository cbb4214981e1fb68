"""The CNN oracle (oracle/cnn.py) pinned against independent references on
the CPU: torch autograd in float64 with rounding off, central differences,
torch's bfloat16 conversion and torch.optim.Adam."""
import numpy as np
import pytest

from oracle import cnn as OC


def _batch(seed, n=6):
    rng = np.random.default_rng(seed)
    x = OC.normalize_patches(rng.normal(0, 1, (n, 121)).astype(np.float32) + np.float32(0.3))
    t = np.stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), (rng.random(n) < 0.6)], 1)
    return x, t.astype(np.float32)


def _params(seed):
    p = OC.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    # non-zero biases so every term of the gradient is exercised
    for k in ("b1", "b2", "b3", "bfc"):
        o, s = OC.OFF[k]
        p[o:o + int(np.prod(s))] = rng.uniform(-0.1, 0.1, int(np.prod(s)))
    return p


def _torch_model(p):
    import torch
    P = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in OC.unpack(p).items()}

    def f(x):
        F = torch.nn.functional
        a = F.relu(F.conv2d(x, P["w1"], P["b1"]))
        a = F.relu(F.conv2d(a, P["w2"], P["b2"]))
        a = F.relu(F.conv2d(a, P["w3"], P["b3"]))
        g = a.mean(dim=(2, 3))
        return g @ P["wfc"].T + P["bfc"]
    return f, P


def test_bf16_matches_torch_conversion():
    import torch
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e-30, 100), rng.normal(0, 1e30, 100)])
    v = v.astype(np.float32)
    # exact ties: the low 16 bits are 0x8000
    ties = (rng.integers(0, 2**31, 200, dtype=np.uint32) & 0xFFFF0000) | 0x8000
    v = np.concatenate([v, ties.view(np.float32)[np.isfinite(ties.view(np.float32))]])
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(OC.bf16(v).view(np.uint32), ref.view(np.uint32))


def test_forward_matches_torch_without_rounding():
    import torch
    p = _params(1)
    x, _ = _batch(1)
    f, _ = _torch_model(p)
    ref = f(torch.tensor(x[:, None], dtype=torch.float64)).detach().numpy()
    got = OC.forward(p, x, rounding=False)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_gradients_match_torch_autograd_without_rounding():
    import torch
    p = _params(2)
    x, t = _batch(2, n=8)
    f, P = _torch_model(p)
    o = f(torch.tensor(x[:, None], dtype=torch.float64))
    tt = torch.tensor(t, dtype=torch.float64)
    lab = tt[:, 2]
    reg = lab * ((o[:, 0] - tt[:, 0]) ** 2 + (o[:, 1] - tt[:, 1]) ** 2)
    bce = torch.nn.functional.binary_cross_entropy_with_logits(o[:, 2], lab, reduction="none")
    L = (reg + bce).mean()
    L.backward()
    ref = np.concatenate([P[k].grad.numpy().reshape(-1) for k, _ in OC.LAYOUT])
    got, lval = OC.gradients(p, x, t, rounding=False)
    assert abs(lval - L.item()) < 1e-12
    np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-13)


def test_gradients_match_central_differences():
    p = _params(3).astype(np.float64)
    x, t = _batch(3, n=4)
    g, _ = OC.gradients(p, x, t, rounding=False)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(np.arange(o, o + int(np.prod(s))), 3, replace=False)
                          for o, s in OC.OFF.values()])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        # forward() takes float32 parameters: evaluate the loss in float64 directly
        num = (_loss64(pp, x, t) - _loss64(pm, x, t)) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, num, g[i])


def _loss64(p, x, t):
    # forward in float64 without the float32 cast of the parameters
    P = {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OC.OFF.items()}
    xx = np.asarray(x, np.float64).reshape(-1, 1, 11, 11)
    a = np.maximum(OC._conv(xx, P["w1"], P["b1"]), 0)
    a = np.maximum(OC._conv(a, P["w2"], P["b2"]), 0)
    a = np.maximum(OC._conv(a, P["w3"], P["b3"]), 0)
    o = a.mean(axis=(2, 3)) @ P["wfc"].T + P["bfc"]
    return OC.loss(o, t)


def test_rounding_changes_results_at_bf16_scale_only():
    p = _params(4)
    x, t = _batch(4, n=16)
    exact = OC.forward(p, x, rounding=False)
    rnd = OC.forward(p, x, rounding=True)
    err = np.abs(rnd - exact).max() / (np.abs(exact).max() + 1e-12)
    assert 0 < err < 0.05
    g0, _ = OC.gradients(p, x, t, rounding=False)
    g1, _ = OC.gradients(p, x, t, rounding=True)
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 0.05


def test_adam_matches_torch():
    import torch
    rng = np.random.default_rng(5)
    p0 = rng.normal(0, 1, 50).astype(np.float32)
    grads = [rng.normal(0, 1, 50).astype(np.float32) for _ in range(6)]
    w = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([w], lr=1e-3)
    p, m, v = p0.copy(), np.zeros(50, np.float32), np.zeros(50, np.float32)
    for t, g in enumerate(grads, 1):
        w.grad = torch.tensor(g)
        opt.step()
        p, m, v = OC.adam_step(p, g, m, v, t)
    np.testing.assert_allclose(p, w.detach().numpy(), rtol=0, atol=2e-7)


def test_patch_gather_and_normalisation():
    proc = np.arange(20 * 30, dtype=np.float32).reshape(20, 30) - 100.0
    raw = OC.gather_patches(proc, [(2, 3), (10, 15)])
    # (2, 3): rows -3..7 and columns -2..8 -> rows 0..2 and columns 0..1 outside
    r = raw[0].reshape(11, 11)
    assert np.all(r[:3] == 0) and np.all(r[:, :2] == 0)
    assert r[5, 5] == proc[2, 3] and r[10, 10] == proc[7, 8]
    x = OC.normalize_patches(raw)
    s = np.abs(raw[1]).max()
    assert s == proc[15, 20]
    assert x[1, 10, 10] == 1.0
    assert x[1, 0, 0] == OC.bf16(np.float32(proc[5, 10]) / np.float32(s))
    assert np.all(OC.normalize_patches(np.zeros((1, 121), np.float32)) == 0)


def test_label_zero_ignores_the_target_offset():
    p = _params(6)
    x, t = _batch(6, n=6)
    t[:, 2] = [1, 0, 1, 0, 0, 1]
    t2 = t.copy()
    t2[t[:, 2] == 0, :2] = np.nan
    g1, l1 = OC.gradients(p, x, t)
    g2, l2 = OC.gradients(p, x, t2)
    assert l1 == l2 and np.array_equal(g1, g2)
