"""Seeded synthetic sequences for the search benchmarks and tests.

Registered frames of a static field: sky level + Gaussian noise, static stars with a
pixel-integrated Gaussian PSF, and linear movers trailed over the exposure, at a fixed
cadence. Movers are given a per-frame matched-filter S/N, so the flux is
F = snr * sigma_noise * sqrt(4 pi) * sigma_psf (the S/N of a Gaussian PSF under a
matched filter in white noise). Frames are built with torch on the requested device.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Mover:
    x0: float  # position at the reference epoch (middle of the sequence), px
    y0: float
    vx_pxh: float
    vy_pxh: float
    flux: float
    inv_dist_au: float = 0.0  # curved track: plus inv_dist_au times the frame's parallax displacement


def diurnal_parallax(t_h, amp_px: float = 9.0, phase: float = 0.4):
    """Parallax displacements at 1 AU (pixels) of a ground observer over a night: the
    diurnal circle's projection, amp_px at most, less its least-squares line in t (as
    set_parallax_model leaves them, src/parallax.cpp:138-155, so that the velocity grid
    absorbs the straight part)."""
    t = np.asarray(t_h, np.float64)
    w = 2.0 * math.pi / 23.934  # sidereal rate, 1/h
    px = amp_px * np.cos(w * t + phase)
    py = 0.45 * amp_px * np.sin(w * t + phase)
    A = np.stack([np.ones_like(t), t], 1)
    for v in (px, py):
        v -= A @ np.linalg.lstsq(A, v, rcond=None)[0]
    return px, py


def _gauss_pixel_weights(c: float, lo: int, hi: int, sigma: float) -> np.ndarray:
    xs = np.arange(lo, hi + 1, dtype=np.float64)
    inv = 1.0 / (sigma * math.sqrt(2.0))
    erf = np.vectorize(math.erf)
    return 0.5 * (erf((xs + 0.5 - c) * inv) - erf((xs - 0.5 - c) * inv))


def _add_source(img: np.ndarray, x0: float, y0: float, flux: float, sigma: float) -> None:
    h, w = img.shape
    r = int(math.ceil(sigma * math.sqrt(2.0 * math.log(max(flux / (2 * math.pi * sigma * sigma) / 0.05, 2.0)))))
    r = min(max(r, 2), 30)
    xc, yc = int(math.floor(x0 + 0.5)), int(math.floor(y0 + 0.5))
    xl, xh = max(0, xc - r), min(w - 1, xc + r)
    yl, yh = max(0, yc - r), min(h - 1, yc + r)
    if xl > xh or yl > yh:
        return
    wx = _gauss_pixel_weights(x0, xl, xh, sigma)
    wy = _gauss_pixel_weights(y0, yl, yh, sigma)
    img[yl:yh + 1, xl:xh + 1] += (flux * np.outer(wy, wx)).astype(np.float32)


def make_sequence(n: int, w: int, h: int, *, sky: float = 1000.0, noise: float = 32.0, nstars: int = 150,
                  star_flux=(1e3, 1e6), psf_sigma: float = 1.5, nmovers: int = 3, mover_snr=(1.5, 1.5),
                  rate_px_min=(0.2, 1.1), cadence_s: float = 60.0, exptime_s: float = 60.0, seed: int = 7,
                  device: str = "cpu", mjd0: float = 61000.0, saturate: float = 0.0, cr_frames: int = 0,
                  cr_hits: int = 50, mover_box=None, parallax=None, mover_inv_dist=(0.0, 0.0)):
    """Returns (frames (n, h, w) float32 torch tensor on `device`, mjd (n,), movers).

    saturate > 0 clips every pixel at that level (a detector's full well, e.g. 65535
    ADU); cr_frames frames (chosen by the seed) get cr_hits cosmic-ray hits of 1-4
    pixels and 800-15000 ADU each. mover_box (x0, x1, y0, y1): where the movers are at
    the reference epoch (default: 24 px inside the frame). parallax (par_x, par_y), one
    per frame: movers then follow curved tracks x + v t + q par, q drawn from
    mover_inv_dist (1/AU)."""
    import torch
    rng = np.random.default_rng(seed)
    static = np.zeros((h, w), np.float32)
    for _ in range(nstars):
        x0, y0 = rng.uniform(-5, w + 5), rng.uniform(-5, h + 5)
        f = math.exp(rng.uniform(math.log(star_flux[0]), math.log(star_flux[1])))
        _add_source(static, x0, y0, f, psf_sigma)
    t_s = np.arange(n) * cadence_s
    mjd = mjd0 + t_s / 86400.0
    t_ref = 0.5 * (mjd[0] + mjd[-1])
    t_h = (mjd - t_ref) * 24.0
    bx0, bx1, by0, by1 = mover_box if mover_box is not None else (24, w - 24, 24, h - 24)
    movers = []
    for _ in range(nmovers):
        snr = rng.uniform(*mover_snr)
        rate = rng.uniform(*rate_px_min) * 60.0  # px/h
        pa = rng.uniform(0.0, 2 * math.pi)
        movers.append(Mover(rng.uniform(bx0, bx1), rng.uniform(by0, by1), rate * math.cos(pa), rate * math.sin(pa),
                            snr * noise * math.sqrt(4 * math.pi) * psf_sigma,
                            float(rng.uniform(*mover_inv_dist)) if parallax is not None else 0.0))
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    base = torch.from_numpy(static).to(device) + sky
    frames = torch.empty((n, h, w), dtype=torch.float32, device=device)
    steps = 8  # sub-positions across the exposure for the trail
    for i in range(n):
        mv = np.zeros((h, w), np.float32)
        for m in movers:
            cx = m.inv_dist_au * parallax[0][i] if parallax is not None else 0.0
            cy = m.inv_dist_au * parallax[1][i] if parallax is not None else 0.0
            for k in range(steps):
                tt = t_h[i] + ((k + 0.5) / steps - 0.5) * exptime_s / 3600.0
                _add_source(mv, m.x0 + m.vx_pxh * tt + cx, m.y0 + m.vy_pxh * tt + cy, m.flux / steps, psf_sigma)
        frames[i] = base + torch.from_numpy(mv).to(device) + noise * torch.randn((h, w), generator=gen, device=device)
    for i in rng.choice(n, size=min(cr_frames, n), replace=False) if cr_frames > 0 else []:
        for _ in range(cr_hits):
            x, y = int(rng.integers(0, w)), int(rng.integers(0, h))
            amp = float(rng.uniform(800.0, 15000.0))
            for _ in range(int(rng.integers(1, 5))):
                frames[i, y, x] += amp
                x = min(w - 1, max(0, x + int(rng.integers(-1, 2))))
                y = min(h - 1, max(0, y + int(rng.integers(-1, 2))))
    if saturate > 0.0:
        frames.clamp_(max=saturate)
    return frames, mjd, movers
