"""Helpers for the GPU tests: device buffers with padded pitches and C-ABI calls.

Every output buffer is allocated with a few extra columns so that the row pitch
differs from the row width (the C-ABI takes pitches in bytes).
"""
from __future__ import annotations

import numpy as np
import torch

import pyrsgm_amd as P

PAD = 7  # extra columns in every pitched device buffer


class Img:
    """A pitched 2-D device buffer."""

    def __init__(self, h, w, dtype, fill=0):
        self.h, self.w = h, w
        self.t = torch.full((h, w + PAD), fill, dtype=dtype, device="cuda")

    @classmethod
    def upload(cls, a: np.ndarray):
        dt = {np.uint8: torch.uint8, np.int16: torch.int16, np.float32: torch.float32}[a.dtype.type]
        im = cls(a.shape[0], a.shape[1], dt)
        im.t[:, :a.shape[1]] = torch.from_numpy(np.ascontiguousarray(a)).cuda()
        return im

    @property
    def ptr(self):
        return self.t.data_ptr()

    @property
    def pitch(self):
        return self.t.stride(0) * self.t.element_size()

    def numpy(self):
        return self.t[:, :self.w].cpu().numpy()


def stream():
    return torch.cuda.current_stream().cuda_stream


def sync():
    torch.cuda.synchronize()


def single_scale(ws, left, right, D, minD=0, block=5, P1=300, P2=2400, uniq=10, disp12=1, cap=63,
                 cost=0, post=3, paths=3, adaptive=0, alpha=3.041, beta=16.137, lower=0.601,
                 upper=1.77, want16=True, want32=True, wantc=True):
    h, w = left.shape
    L, R = Img.upload(left), Img.upload(right)
    d16 = Img(h, w, torch.int16, 1234) if want16 else None
    d32 = Img(h, w, torch.float32, -7.0) if want32 else None
    cf = Img(h, w, torch.float32, -7.0) if wantc else None
    rc = P.compute_single_scale(
        ws, L.ptr, R.ptr, d16.ptr if d16 else None, d32.ptr if d32 else None,
        cf.ptr if cf else None, L.pitch, R.pitch, d16.pitch if d16 else 0,
        d32.pitch if d32 else 0, cf.pitch if cf else 0, w, h, D, minD, block, P1, P2, uniq,
        disp12, cap, cost, post, paths, adaptive, alpha, beta, lower, upper, stream())
    sync()
    return rc, (d16.numpy() if d16 else None), (d32.numpy() if d32 else None), (cf.numpy() if cf else None)


def level(ws, left, right, D, maxD, parent_disp=None, parent_conf=None, minD=0, block=5, P1=300,
          P2=2400, uniq=10, disp12=1, cap=63, cost=0, post=1, paths=3, refine=False,
          min_conf=0.05, adaptive=0, alpha=3.041, beta=16.137, lower=0.601, upper=1.77,
          want16=True, want32=True, wantc=True):
    h, w = left.shape
    L, R = Img.upload(left), Img.upload(right)
    pd = Img.upload(parent_disp) if parent_disp is not None else None
    pc = Img.upload(parent_conf) if parent_conf is not None else None
    d16 = Img(h, w, torch.int16, 1234) if want16 else None
    d32 = Img(h, w, torch.float32, -7.0) if want32 else None
    cf = Img(h, w, torch.float32, -7.0) if wantc else None
    ph, pw = (parent_disp.shape if parent_disp is not None else (0, 0))
    rc = P.compute_level(
        ws, L.ptr, R.ptr, pd.ptr if pd else None, pc.ptr if pc else None,
        d16.ptr if d16 else None, d32.ptr if d32 else None, cf.ptr if cf else None,
        L.pitch, R.pitch, pd.pitch if pd else 0, pc.pitch if pc else 0,
        d16.pitch if d16 else 0, d32.pitch if d32 else 0, cf.pitch if cf else 0,
        w, h, pw, ph, D, minD, maxD, block, P1, P2, uniq, disp12, cap, cost, post, paths,
        int(refine), min_conf, adaptive, alpha, beta, lower, upper, stream())
    sync()
    return rc, (d16.numpy() if d16 else None), (d32.numpy() if d32 else None), (cf.numpy() if cf else None)


def matcher(left, right, params, want16=True, want32=True, wantc=True):
    h, w = left.shape
    m = P.StereoMatcher(w, h, params)
    L, R = Img.upload(left), Img.upload(right)
    d16 = Img(h, w, torch.int16, 1234) if want16 else None
    d32 = Img(h, w, torch.float32, -7.0) if want32 else None
    cf = Img(h, w, torch.float32, -7.0) if wantc else None
    m.compute_ptr(L.ptr, L.pitch, R.ptr, R.pitch, d16.ptr if d16 else None,
                  d16.pitch if d16 else 0, d32.ptr if d32 else None, d32.pitch if d32 else 0,
                  cf.ptr if cf else None, cf.pitch if cf else 0, stream())
    sync()
    out = ((d16.numpy() if d16 else None), (d32.numpy() if d32 else None),
           (cf.numpy() if cf else None))
    m.close()
    return out


def assert_same(name, got, want):
    if want is None:
        return
    assert got is not None, name
    if got.dtype == np.float32:
        bad = ~((got == want) | (np.isnan(got) & np.isnan(want)))
    else:
        bad = got != want
    n = int(bad.sum())
    if n:
        ys, xs = np.nonzero(bad)
        detail = ", ".join(f"({y},{x}): got {got[y, x]} want {want[y, x]}"
                           for y, x in list(zip(ys, xs))[:8])
        raise AssertionError(f"{name}: {n} of {got.size} differ; first: {detail}")
