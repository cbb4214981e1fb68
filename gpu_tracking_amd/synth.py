"""Seeded synthetic particle videos for tests and benchmarks.

Gaussian spots sampled at pixel centres, Brownian motion reflected at a
border band, optional blinking, a constant or linear-ramp background and
Poisson shot noise (uint16) or Gaussian read noise (float32).  The named
workloads follow the benchmark configurations of this project:

  small    512x512 u16, 200 particles, sigma 1.5, A ~ lognormal(ln 800, 0.4),
           D = 0.5, background 100 + Poisson
  full     1024x1024 u16, 1000 particles, sigma ~ U[1.2, 2.0],
           A ~ lognormal(ln 600, 0.5), D = 1.0, 5% blinking,
           background ramp 80..140 along x + Poisson
  log      2048x2048 f32, 3000 particles, sigma ~ U[1, 5], A ~ U[200, 2000],
           background 50 + N(0, 5^2)

``video_numpy`` runs on the host (numpy); ``video_torch`` renders the same
model on the GPU with torch (a different random stream: use it when the
frames only need to be statistically alike, e.g. for throughput).
"""
from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np


@dataclass(frozen=True)
class Spec:
    nrows: int
    ncols: int
    n_particles: int
    sigma: tuple = (1.5, 1.5)          # uniform range per particle
    amp: tuple = ("lognormal", np.log(800.0), 0.4)
    D: float = 0.5                     # px^2 / frame
    blink: float = 0.0
    background: tuple = (100.0, 100.0)  # value at x = 0 and x = ncols-1
    noise: str = "poisson"             # "poisson" (uint16) or "gauss:<std>" (float32)
    edge: float = 10.0
    min_dist: float = 10.0


SPECS = {
    "small": Spec(512, 512, 200),
    "full": Spec(1024, 1024, 1000, sigma=(1.2, 2.0), amp=("lognormal", np.log(600.0), 0.5),
                 D=1.0, blink=0.05, background=(80.0, 140.0)),
    "log": Spec(2048, 2048, 3000, sigma=(1.0, 5.0), amp=("uniform", 200.0, 2000.0), D=0.5,
                background=(50.0, 50.0), noise="gauss:5"),
}


def spec(name: str, **overrides) -> Spec:
    return replace(SPECS[name], **overrides)


def _initial(rng, s: Spec):
    pts = []
    lo_y, hi_y = s.edge, s.nrows - 1 - s.edge
    lo_x, hi_x = s.edge, s.ncols - 1 - s.edge
    tries = 0
    while len(pts) < s.n_particles and tries < 200 * s.n_particles:
        tries += 1
        c = np.array([rng.uniform(lo_y, hi_y), rng.uniform(lo_x, hi_x)])
        if pts:
            d2 = np.sum((np.asarray(pts) - c) ** 2, axis=1)
            if np.any(d2 < s.min_dist ** 2):
                continue
        pts.append(c)
    pos = np.asarray(pts, np.float64).reshape(-1, 2)
    n = len(pos)
    sig = rng.uniform(s.sigma[0], s.sigma[1], n) if s.sigma[0] != s.sigma[1] else np.full(n, s.sigma[0])
    if s.amp[0] == "lognormal":
        amp = rng.lognormal(s.amp[1], s.amp[2], n)
    else:
        amp = rng.uniform(s.amp[1], s.amp[2], n)
    return pos, sig, amp


def _reflect(v, lo, hi):
    v = np.where(v < lo, 2 * lo - v, v)
    v = np.where(v > hi, 2 * hi - v, v)
    return np.clip(v, lo, hi)


def trajectories(s: Spec, nframes: int, seed: int):
    """Per-frame particle positions (T, n, 2), sigmas, amplitudes, visibility (T, n)."""
    rng = np.random.default_rng(seed)
    pos, sig, amp = _initial(rng, s)
    n = len(pos)
    out = np.empty((nframes, n, 2))
    vis = np.ones((nframes, n), bool)
    step = np.sqrt(2.0 * s.D)
    for t in range(nframes):
        if t:
            pos = pos + rng.normal(0.0, step, pos.shape)
            pos[:, 0] = _reflect(pos[:, 0], s.edge, s.nrows - 1 - s.edge)
            pos[:, 1] = _reflect(pos[:, 1], s.edge, s.ncols - 1 - s.edge)
        out[t] = pos
        if s.blink > 0:
            vis[t] = rng.random(n) >= s.blink
    return out, sig, amp, vis, rng


def _render(pos, sig, amp, vis_t, s: Spec) -> np.ndarray:
    H, W = s.nrows, s.ncols
    R = int(np.ceil(4 * max(s.sigma)))
    off = np.arange(-R, R + 1)
    p = pos[vis_t]
    sg = sig[vis_t]
    am = amp[vis_t]
    cy = np.rint(p[:, 0]).astype(np.int64)
    cx = np.rint(p[:, 1]).astype(np.int64)
    yy = cy[:, None, None] + off[None, :, None]
    xx = cx[:, None, None] + off[None, None, :]
    dy = yy - p[:, 0, None, None]
    dx = xx - p[:, 1, None, None]
    val = am[:, None, None] * np.exp(-(dy * dy + dx * dx) / (2.0 * sg[:, None, None] ** 2))
    ok = (yy >= 0) & (yy < H) & (xx >= 0) & (xx < W)
    ok = np.broadcast_to(ok, val.shape)
    idx = (np.broadcast_to(yy, val.shape) * W + np.broadcast_to(xx, val.shape))[ok]
    img = np.bincount(idx, weights=val[ok], minlength=H * W).reshape(H, W)
    bg = np.linspace(s.background[0], s.background[1], W)[None, :]
    return img + bg


def video_numpy(s: Spec, nframes: int, seed: int) -> np.ndarray:
    """(T, H, W) uint16 (Poisson noise) or float32 (Gaussian read noise)."""
    traj, sig, amp, vis, rng = trajectories(s, nframes, seed)
    gauss = s.noise.startswith("gauss")
    out = np.empty((nframes, s.nrows, s.ncols), np.float32 if gauss else np.uint16)
    for t in range(nframes):
        lam = _render(traj[t], sig, amp, vis[t], s)
        if gauss:
            std = float(s.noise.split(":")[1])
            out[t] = (lam + rng.normal(0.0, std, lam.shape)).astype(np.float32)
        else:
            out[t] = np.clip(rng.poisson(lam), 0, 65535).astype(np.uint16)
    return out


def video_torch(s: Spec, nframes: int, seed: int, device="cuda", batch: int = 64):
    """The same model rendered on the GPU: (T, H, W) int16 holding u16 values
    (Poisson) or float32 (Gaussian).  Trajectories come from numpy."""
    import torch
    traj, sig, amp, vis, _ = trajectories(s, nframes, seed)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    H, W = s.nrows, s.ncols
    gauss = s.noise.startswith("gauss")
    out = torch.empty((nframes, H, W), dtype=torch.float32 if gauss else torch.int16,
                      device=device)
    R = int(np.ceil(4 * max(s.sigma)))
    off = torch.arange(-R, R + 1, device=device, dtype=torch.float64)
    bg = torch.linspace(s.background[0], s.background[1], W, device=device, dtype=torch.float64)
    sig_t = torch.as_tensor(sig, device=device)
    amp_t = torch.as_tensor(amp, device=device)
    for t0 in range(0, nframes, batch):
        t1 = min(nframes, t0 + batch)
        nb = t1 - t0
        p = torch.as_tensor(traj[t0:t1], device=device)            # (nb, n, 2)
        v = torch.as_tensor(vis[t0:t1], device=device)              # (nb, n)
        cy = torch.round(p[..., 0]).long()
        cx = torch.round(p[..., 1]).long()
        yy = cy[..., None, None] + off.long()[None, None, :, None]
        xx = cx[..., None, None] + off.long()[None, None, None, :]
        dy = yy.double() - p[..., 0, None, None]
        dx = xx.double() - p[..., 1, None, None]
        val = amp_t[None, :, None, None] * torch.exp(
            -(dy * dy + dx * dx) / (2.0 * sig_t[None, :, None, None] ** 2))
        ok = (yy >= 0) & (yy < H) & (xx >= 0) & (xx < W) & v[..., None, None]
        fidx = torch.arange(nb, device=device)[:, None, None, None]
        flat = (fidx * (H * W) + yy * W + xx)[ok]
        img = torch.zeros(nb * H * W, dtype=torch.float64, device=device)
        img.index_add_(0, flat, val[ok])
        lam = img.view(nb, H, W) + bg[None, None, :]
        if gauss:
            std = float(s.noise.split(":")[1])
            noise = torch.randn(lam.shape, generator=g, device=device, dtype=torch.float64) * std
            out[t0:t1] = (lam + noise).float()
        else:
            cnt = torch.poisson(lam.float(), generator=g)
            out[t0:t1] = torch.clamp(cnt, 0, 65535).to(torch.int32).to(torch.int16)
    return out
