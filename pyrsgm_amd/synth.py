"""Seeded synthetic rectified stereo pairs (the workload generator).

A pair is built as follows, all from ``numpy.random.default_rng(seed)``:

* left image: uniform uint8 noise, separable Gaussian blur (sigma 1.5, radius 5,
  REFLECT_101 borders, round half up), then three constant (textureless)
  rectangles, each 10-30 % of the image width/height, at random places and
  grey levels;
* ground truth: ``regions`` piecewise-planar patches (a grid of random vertical
  and horizontal cuts), each ``d = d0 + bx*(x-cx) + by*(y-cy)`` clipped to
  ``[0, max_disp]``;
* right image: forward warp ``xr = round(x - d)`` of the left image, keeping the
  nearer surface (larger d) where several pixels land on one column, holes
  filled from the nearest written pixel to their right (left at the border),
  plus N(0, 2) noise, rounded and clamped to uint8.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Pair:
    left: np.ndarray    # (H, W) uint8
    right: np.ndarray   # (H, W) uint8
    gt: np.ndarray      # (H, W) float32 disparity of the left view
    seed: int


def _reflect101(idx: np.ndarray, n: int) -> np.ndarray:
    idx = np.abs(idx)
    return np.where(idx >= n, 2 * n - 2 - idx, idx)


def _blur(img: np.ndarray, sigma: float = 1.5, radius: int = 5, chunk: int = 32) -> np.ndarray:
    # Processed in slabs of `chunk` rows (identical arithmetic, cache-sized temporaries).
    t = np.arange(-radius, radius + 1, dtype=np.float64)
    k = np.exp(-0.5 * (t / sigma) ** 2)
    k /= k.sum()
    h, w = img.shape
    f = img.astype(np.float64)
    cols = _reflect101(np.arange(w)[:, None] + t[None, :].astype(int), w)
    g = np.empty((h, w))
    for y0 in range(0, h, chunk):
        g[y0:y0 + chunk] = (f[y0:y0 + chunk][:, cols] * k).sum(axis=2)
    rows = _reflect101(np.arange(h)[:, None] + t[None, :].astype(int), h)
    out = np.empty((h, w))
    for y0 in range(0, h, chunk):
        out[y0:y0 + chunk] = (g[rows[y0:y0 + chunk], :] * k[None, :, None]).sum(axis=1)
    return np.clip(np.floor(out + 0.5), 0, 255).astype(np.uint8)


def _planes(rng, w: int, h: int, max_disp: int, regions: int) -> np.ndarray:
    ny = 2 if regions >= 4 else 1
    nx = max(1, regions // ny)
    xc = np.sort(rng.integers(w // 8, w - w // 8, size=nx - 1)) if nx > 1 else np.array([], int)
    yc = np.sort(rng.integers(h // 8, h - h // 8, size=ny - 1)) if ny > 1 else np.array([], int)
    xe = np.concatenate([[0], xc, [w]])
    ye = np.concatenate([[0], yc, [h]])
    gt = np.zeros((h, w), np.float64)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    for iy in range(ny):
        for ix in range(nx):
            y0, y1, x0, x1 = ye[iy], ye[iy + 1], xe[ix], xe[ix + 1]
            d0 = rng.uniform(0.1, 0.9) * max_disp
            bx = rng.uniform(-0.25, 0.25) * max_disp / w
            by = rng.uniform(-0.25, 0.25) * max_disp / h
            cx, cy = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
            sl = (slice(y0, y1), slice(x0, x1))
            gt[sl] = d0 + bx * (xx[sl] - cx) + by * (yy[sl] - cy)
    return np.clip(gt, 0, max_disp)


def _warp(left: np.ndarray, gt: np.ndarray) -> np.ndarray:
    h, w = left.shape
    xs = np.broadcast_to(np.arange(w), (h, w))
    ys = np.broadcast_to(np.arange(h)[:, None], (h, w))
    xr = np.rint(xs - gt).astype(np.int64)
    ok = (xr >= 0) & (xr < w)
    tgt = (ys * w + xr)[ok]
    dsrc = gt[ok]
    src = (ys * w + xs)[ok]
    zbuf = np.full(h * w, -1.0)
    np.maximum.at(zbuf, tgt, dsrc)
    win = dsrc == zbuf[tgt]
    best_src = np.full(h * w, -1, np.int64)
    np.maximum.at(best_src, tgt[win], src[win])
    right = np.zeros(h * w, np.float64)
    filled = best_src >= 0
    right[filled] = left.reshape(-1)[best_src[filled]]
    right = right.reshape(h, w)
    filled = filled.reshape(h, w)
    # Holes take the nearest written value to their right (left at the border).
    idx = np.where(filled, np.arange(w)[None, :], w)
    nxt = np.minimum.accumulate(idx[:, ::-1], axis=1)[:, ::-1]
    idx_l = np.where(filled, np.arange(w)[None, :], -1)
    prv = np.maximum.accumulate(idx_l, axis=1)
    src_col = np.where(nxt < w, nxt, np.maximum(prv, 0))
    right = np.take_along_axis(right, src_col, axis=1)
    return right


def make_pair(width: int, height: int, max_disp: int, seed: int, regions: int = 4) -> Pair:
    rng = np.random.default_rng(seed)
    left = _blur(rng.integers(0, 256, size=(height, width), dtype=np.uint8))
    for _ in range(3):
        rw = int(rng.uniform(0.1, 0.3) * width)
        rh = int(rng.uniform(0.1, 0.3) * height)
        x0 = int(rng.integers(0, max(1, width - rw)))
        y0 = int(rng.integers(0, max(1, height - rh)))
        left[y0:y0 + rh, x0:x0 + rw] = int(rng.integers(0, 256))
    gt = _planes(rng, width, height, max_disp, regions)
    right = _warp(left, gt) + rng.normal(0.0, 2.0, size=(height, width))
    right = np.clip(np.floor(right + 0.5), 0, 255).astype(np.uint8)
    return Pair(left, right, gt.astype(np.float32), seed)


# Workloads named by what they are (sizes, range, cost, pyramid depth).
UNEVEN_SIZES = ((640, 480), (960, 540), (1280, 720), (1920, 1080))
SERVING_SIZES = ((640, 480, 64, 0), (1280, 720, 128, 1), (1920, 1080, 256, 2),
                 (2560, 1440, 256, 2), (3840, 2160, 512, 3))  # (w, h, D, pyramid_levels)


def uneven_spec(i: int, seed0: int = 2000) -> dict:
    """Pair i of the uneven batch: size from UNEVEN_SIZES, D from {64, 128,
    256}, pyramid_levels 0-3 (one to four levels in all) and BT or Census,
    each drawn uniformly from the pair's own seed; the 3x3 median is on for
    the even-numbered half.  Its images come from make_pair with the same
    seed and disparities up to D - 1."""
    rng = np.random.default_rng(seed0 + i)
    w, h = UNEVEN_SIZES[int(rng.integers(len(UNEVEN_SIZES)))]
    D = (64, 128, 256)[int(rng.integers(3))]
    levels = int(rng.integers(4))
    cost = ("bt", "census")[int(rng.integers(2))]
    return {"w": w, "h": h, "D": D, "levels": levels, "cost": cost, "median": i % 2 == 0,
            "seed": seed0 + i}


def serving_spec(j: int, seed0: int = 3000) -> dict:
    """Request j of the serving stream: a size from SERVING_SIZES (with its D
    and pyramid depth) and BT or Census, drawn from the request's own seed."""
    rng = np.random.default_rng(seed0 + j)
    w, h, D, levels = SERVING_SIZES[int(rng.integers(len(SERVING_SIZES)))]
    cost = ("bt", "census")[int(rng.integers(2))]
    return {"w": w, "h": h, "D": D, "levels": levels, "cost": cost, "median": True,
            "seed": seed0 + j}


def spec_pair(spec: dict) -> Pair:
    return make_pair(spec["w"], spec["h"], spec["D"] - 1, seed=spec["seed"])


def spec_pairs(specs, workers: int = 8) -> list:
    """make_pair for many specs on a thread pool (numpy releases the GIL in
    the heavy array operations)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        return list(ex.map(spec_pair, specs))


def config_pair(name: str, index: int = 0) -> Pair:
    if name == "640x480_d64_bt_single":
        return make_pair(640, 480, 63, seed=0, regions=4)
    if name == "1920x1080_d256_census_l2":
        return make_pair(1920, 1080, 255, seed=1, regions=8)
    if name == "960x540_d128_census_l2":
        return make_pair(960, 540, 127, seed=1000 + index, regions=4)
    raise KeyError(name)
