"""Bit-exact parity of the HIP path with the CPU oracle, through the C-ABI.

Integer outputs (disparity*16) must be identical; the float outputs (disparity
in pixels, confidence) are compared exactly too, which is stricter than the
1e-6 absolute tolerance the engine promises.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pyrsgm_amd import synth
from tests import gpu_util as U

pytestmark = pytest.mark.gpu

import pyrsgm_amd as P


@pytest.fixture(scope="module")
def ws():
    w = P.create_workspace()
    yield w
    P.destroy_workspace(w)


def pattern(w, h, d, right):
    xs = np.arange(w)[None, :] + (d if right else 0)
    ys = np.arange(h)[:, None]
    return ((xs * 37 + ys * 17 + ((xs // 7) ^ (ys // 5)) * 29) & 0xFF).astype(np.uint8)


def pair(w, h, maxd, seed):
    p = synth.make_pair(w, h, maxd, seed=seed)
    return p.left, p.right


# --------------------------------------------------------------- pyrDown
# Sizes from 643 x 301 up have interior blocks (no reflection) beside border
# blocks; off = 1 starts the source rows one byte past a 4-byte boundary.
@pytest.mark.parametrize("w,h,off", [(64, 48, 0), (64, 33, 0), (97, 48, 0), (97, 33, 0), (2, 2, 0),
                                     (3, 5, 0), (643, 301, 0), (1920, 1080, 0), (643, 301, 1)])
def test_pyrdown_kat(w, h, off):
    xs, ys = np.meshgrid(np.arange(w + off), np.arange(h))
    full = ((xs * 13 + ys * 29 + (xs * ys) // 7) & 0xFF).astype(np.uint8)
    src = full[:, off:]
    want = O.pyrdown(np.ascontiguousarray(src))
    torch = __import__("torch")
    pitch = (w + off + 63) // 64 * 64  # 4-byte aligned rows
    buf = np.zeros((h, pitch), np.uint8)
    buf[:, :w + off] = full
    S = torch.from_numpy(buf).cuda()
    dh, dw = want.shape
    Dd = U.Img(dh, dw, torch.uint8, 77)
    print("case: pyrdown", w, h, "offset", off, flush=True)
    assert P.pyrdown_u8(S.data_ptr() + off, pitch, w, h, Dd.ptr, Dd.pitch, dw, dh, U.stream()) == 0
    U.sync()
    U.assert_same("pyrdown", Dd.numpy(), want)


def test_pyrdown_rejects_bad_sizes():
    import torch
    S = U.Img(10, 10, torch.uint8)
    Dd = U.Img(5, 5, torch.uint8)
    assert P.pyrdown_u8(S.ptr, S.pitch, 10, 10, Dd.ptr, Dd.pitch, 4, 5, 0) == -1
    assert P.pyrdown_u8(S.ptr, S.pitch, 1, 10, Dd.ptr, Dd.pitch, 1, 5, 0) == -1
    assert P.pyrdown_u8(None, S.pitch, 10, 10, Dd.ptr, Dd.pitch, 5, 5, 0) == -1


# ---------------------------------------------------------- single scale
SINGLE = [
    # (w, h, D, minD, cost, post, paths, adaptive, block)
    (96, 64, 16, 0, 0, 3, 3, 0, 5),
    (96, 64, 32, 0, 1, 3, 3, 0, 5),
    (131, 47, 48, 0, 0, 3, 3, 1, 5),
    (131, 47, 64, 0, 1, 1, 3, 1, 5),
    (160, 90, 64, 0, 0, 2, 2, 0, 3),
    (160, 90, 128, 0, 1, 0, 1, 0, 5),
    (200, 37, 64, 5, 0, 3, 3, 0, 7),
    (200, 37, 32, -6, 1, 3, 3, 1, 5),
    (300, 20, 256, 0, 0, 3, 3, 0, 5),
    (70, 33, 80, 0, 1, 3, 3, 0, 5),
]


@pytest.mark.parametrize("w,h,D,minD,cost,post,paths,adaptive,block", SINGLE)
def test_single_scale_parity(ws, w, h, D, minD, cost, post, paths, adaptive, block):
    left, right = pair(w, h, max(4, D - 8), seed=w * 7 + D)
    P1, P2 = (8, 32) if cost else (300, 2400)
    kw = dict(minD=minD, block=block, P1=P1, P2=P2, uniq=10, disp12=1, cap=63, cost=cost,
              post=post, paths=paths, adaptive=adaptive)
    rc, d16, d32, cf = U.single_scale(ws, left, right, D, **kw)
    want = O.compute_single_scale(left, right, D, **kw)
    assert rc == 0 and want.rc == 0
    U.assert_same("disp16", d16, want.disp16)
    U.assert_same("disp32", d32, want.disp32)
    U.assert_same("conf", cf, want.conf)


@pytest.mark.parametrize("cap", [95, 127, 128, 200])
def test_single_scale_bt_pre_filter_cap(ws, cap):
    """The stored BT gradient is a byte (sgm_kernels.cu:117): with preFilterCap
    128 and above it wraps, and the costs follow the wrapped values."""
    left, right = pair(150, 60, 50, seed=cap)
    kw = dict(block=5, P1=300, P2=2400, cap=cap, cost=0, post=3, paths=3, adaptive=0)
    rc, d16, d32, cf = U.single_scale(ws, left, right, 64, **kw)
    want = O.compute_single_scale(left, right, 64, **kw)
    assert rc == 0 and want.rc == 0
    U.assert_same(f"cap {cap} disp16", d16, want.disp16)
    U.assert_same(f"cap {cap} conf", cf, want.conf)


def test_single_scale_output_subsets(ws):
    left, right = pair(96, 40, 40, seed=3)
    for w16, w32, wc in [(True, False, False), (False, True, True), (False, False, True),
                         (False, True, False)]:
        rc, d16, d32, cf = U.single_scale(ws, left, right, 64, want16=w16, want32=w32, wantc=wc)
        want = O.compute_single_scale(left, right, 64, want16=w16, want32=w32, wantc=wc)
        assert rc == 0
        U.assert_same("disp16", d16, want.disp16)
        U.assert_same("disp32", d32, want.disp32)
        U.assert_same("conf", cf, want.conf)


# -------------------------------------------------------------- pyramid
COARSE = [
    # (w, h, D, maxD, cost, paths, adaptive, block)
    (120, 68, 16, 15, 1, 3, 0, 3),
    (120, 68, 32, 31, 0, 3, 1, 3),
    (240, 135, 64, 63, 1, 3, 1, 3),
    (160, 64, 128, 127, 1, 2, 0, 3),
    (100, 50, 64, 40, 0, 1, 0, 5),
]


@pytest.mark.parametrize("w,h,D,maxD,cost,paths,adaptive,block", COARSE)
def test_coarse_level_parity(ws, w, h, D, maxD, cost, paths, adaptive, block):
    left, right = pair(w, h, maxD, seed=w + D)
    P1, P2 = (8, 32) if cost else (300, 2400)
    kw = dict(block=block, P1=P1, P2=P2, cost=cost, post=1, paths=paths, adaptive=adaptive)
    rc, d16, d32, cf = U.level(ws, left, right, D, maxD, **kw)
    want = O.compute_level(left, right, D, maxD, **kw)
    assert rc == 0 and want.rc == 0
    U.assert_same("disp16", d16, want.disp16)
    U.assert_same("disp32", d32, want.disp32)
    U.assert_same("conf", cf, want.conf)


REFINE = [
    # (w, h, D, maxD, cost, paths, adaptive, post)
    (240, 136, 16, 127, 1, 3, 1, 1),
    (241, 135, 16, 100, 0, 3, 0, 3),
    (320, 90, 16, 255, 1, 2, 0, 3),
    (200, 60, 16, 63, 0, 1, 1, 1),
    (200, 60, 32, 127, 1, 3, 0, 1),
]


@pytest.mark.parametrize("w,h,D,maxD,cost,paths,adaptive,post", REFINE)
def test_refine_level_parity(ws, w, h, D, maxD, cost, paths, adaptive, post):
    left, right = pair(w, h, maxD, seed=w * 3 + h)
    P1, P2 = (8, 32) if cost else (300, 2400)
    pl, pr = O.pyrdown(left), O.pyrdown(right)
    pw, ph = pl.shape[1], pl.shape[0]
    pmax = min(maxD >> 1, pw - 1)
    pD = max(16, (pmax + 16) // 16 * 16)
    parent = O.compute_level(pl, pr, pD, pmax, block=3, P1=P1, P2=P2, cost=cost, post=1,
                             want32=False)
    kw = dict(block=5 if cost == 0 else 3, P1=P1, P2=P2, cost=cost, post=post, paths=paths,
              adaptive=adaptive, refine=True, min_conf=0.05)
    rc, d16, d32, cf = U.level(ws, left, right, D, maxD, parent.disp16, parent.conf, **kw)
    want = O.compute_level(left, right, D, maxD, parent.disp16, parent.conf, **kw)
    assert rc == 0 and want.rc == 0
    U.assert_same("disp16", d16, want.disp16)
    U.assert_same("disp32", d32, want.disp32)
    U.assert_same("conf", cf, want.conf)


@pytest.mark.parametrize("refine", [False, True])
def test_pyramid_exact_form_large_p2(ws, refine):
    """Cmax + 2*P2' >= SHRT_MAX selects the path variant that keeps the
    reference's explicit SHRT_MAX tests; it must agree with the oracle too."""
    w, h = 200, 72
    left, right = pair(w, h, 60, seed=77)
    kw = dict(block=5, P1=300, P2=16000, cost=0, post=1, paths=3, adaptive=int(refine))
    if refine:
        pl, pr = O.pyrdown(left), O.pyrdown(right)
        parent = O.compute_level(pl, pr, 32, 31, block=3, P1=300, P2=16000, cost=0, post=1,
                                 want32=False)
        kw.update(refine=True, min_conf=0.05)
        args = (16, 63, parent.disp16, parent.conf)
    else:
        args = (64, 63)
    rc, d16, d32, cf = U.level(ws, left, right, *args, **kw)
    want = O.compute_level(left, right, *args, **kw)
    assert rc == 0 and want.rc == 0
    U.assert_same("disp16", d16, want.disp16)
    U.assert_same("disp32", d32, want.disp32)
    U.assert_same("conf", cf, want.conf)


def test_rejects_negative_penalties(ws):
    left, right = pair(64, 32, 10, seed=5)
    rc, *_ = U.single_scale(ws, left, right, 16, P1=-1, P2=10)
    assert rc == -1
    rc, *_ = U.level(ws, left, right, 16, 15, P1=2, P2=-3)
    assert rc == -1


@pytest.mark.parametrize("block,cap", [(1, 63), (3, 63), (5, 63), (7, 63), (5, 95), (5, 96),
                                       (3, 200)])
def test_bt_refine_window_sizes(ws, block, cap):
    """BT refinement costs: 3x3 and 5x5 windows take the packed interior form
    (packed channel sums up to preFilterCap 95, clamped per position above),
    1x1 and 7x7 the generic loop; border columns always the generic loop.
    preFilterCap 200 makes the stored gradient wrap as a byte."""
    w, h, maxD = 203, 77, 100
    left, right = pair(w, h, maxD, seed=block * 17 + 5)
    pl, pr = O.pyrdown(left), O.pyrdown(right)
    pmax = min(maxD >> 1, pl.shape[1] - 1)
    pD = max(16, (pmax + 16) // 16 * 16)
    parent = O.compute_level(pl, pr, pD, pmax, block=3, P1=300, P2=2400, cost=0, post=1,
                             want32=False, cap=cap)
    for D in (16, 32):
        kw = dict(block=block, P1=300, P2=2400, cost=0, post=3, paths=3, adaptive=1,
                  refine=True, min_conf=0.05, cap=cap)
        rc, d16, d32, cf = U.level(ws, left, right, D, maxD, parent.disp16, parent.conf, **kw)
        want = O.compute_level(left, right, D, maxD, parent.disp16, parent.conf, **kw)
        assert rc == 0 and want.rc == 0
        U.assert_same(f"block {block} D {D} disp16", d16, want.disp16)
        U.assert_same(f"block {block} D {D} conf", cf, want.conf)


def test_bt_refine_sparse_parent(ws):
    """BT refinement where whole 64-pixel row segments have no valid parent
    (every label SHRT_MAX), next to segments with a steep disparity ramp."""
    w, h = 300, 40
    left, right = pair(w, h, 120, seed=99)
    pw, ph = (w + 1) // 2, (h + 1) // 2
    xs = np.arange(pw)[None, :]
    pdisp = np.where((xs // 40) % 2 == 0, -16, (xs * 16 * 3) // 8).astype(np.int16)
    pdisp = np.repeat(pdisp, ph, axis=0)
    pconf = np.full((ph, pw), 0.5, np.float32)
    kw = dict(block=5, P1=300, P2=2400, cost=0, post=3, paths=3, refine=True, min_conf=0.05,
              adaptive=1)
    rc, d16, d32, cf = U.level(ws, left, right, 16, 150, pdisp, pconf, **kw)
    want = O.compute_level(left, right, 16, 150, pdisp, pconf, **kw)
    assert rc == 0 and want.rc == 0
    U.assert_same("disp16", d16, want.disp16)
    U.assert_same("conf", cf, want.conf)


def test_refine_varying_base(ws):
    """Parent disparity alternating 2 / 30 px (reference pyramid_smoke.cpp:115-134)."""
    left, right = pattern(320, 128, 16, False), pattern(320, 128, 16, True)
    pdisp = np.where((np.arange(160)[None, :] & 1) == 1, 30 * 16, 2 * 16).astype(np.int16)
    pdisp = np.repeat(pdisp, 64, axis=0)
    pconf = np.ones((64, 160), np.float32)
    for paths in (1, 3):
        kw = dict(block=3, P1=24, P2=96, uniq=0, disp12=1, cost=0, post=1, paths=paths,
                  refine=True, min_conf=0.0, adaptive=1, alpha=2.0, beta=14.0, lower=0.75,
                  upper=2.0)
        rc, d16, d32, cf = U.level(ws, left, right, 16, 255, pdisp, pconf, **kw)
        want = O.compute_level(left, right, 16, 255, pdisp, pconf, **kw)
        assert rc == 0
        U.assert_same("disp16", d16, want.disp16)
        U.assert_same("conf", cf, want.conf)


# -------------------------------------------------------------- matcher
MATCH = [
    # (w, h, D, levels, cost, adaptive, median)
    (200, 120, 64, 0, 0, 0, 1),
    (200, 120, 64, 1, 1, 0, 1),
    (256, 144, 128, 2, 1, 1, 1),
    (256, 144, 128, 2, 0, 0, 0),
    (321, 161, 64, 3, 1, 1, 1),
    (480, 270, 256, 2, 1, 1, 1),
]


@pytest.mark.parametrize("w,h,D,levels,cost,adaptive,median", MATCH)
def test_matcher_parity(w, h, D, levels, cost, adaptive, median):
    left, right = pair(w, h, D - 1, seed=w + levels)
    p = P.Params(num_disparities=D, pyramid_levels=levels, cost_mode=P.CostMode(cost),
                 p1=8 if cost else 300, p2=32 if cost else 2400, adaptive_p2=bool(adaptive),
                 median_filter=bool(median))
    v = P.validate_params(p)
    d16, d32, cf = U.matcher(left, right, v)
    want = O.matcher_compute(left, right, v)
    assert want.rc == 0
    U.assert_same("disp16", d16, want.disp16)
    U.assert_same("disp32", d32, want.disp32)
    U.assert_same("conf", cf, want.conf)


@pytest.mark.parametrize("kind", ["single", "coarse", "refine"])
def test_census_16bit_volumes(ws, kind):
    """Census with P2 = 300: Cmax + max(P2') > 254, so the volumes stay 16-bit
    (the byte form is only used when every path cost fits in 8 bits)."""
    w, h = 160, 72
    left, right = pair(w, h, 60, seed=91)
    if kind == "single":
        kw = dict(block=5, P1=20, P2=300, cost=1, post=3, paths=3, adaptive=1)
        rc, d16, d32, cf = U.single_scale(ws, left, right, 64, **kw)
        want = O.compute_single_scale(left, right, 64, **kw)
    elif kind == "coarse":
        kw = dict(block=3, P1=20, P2=300, cost=1, post=1, paths=3, adaptive=1)
        rc, d16, d32, cf = U.level(ws, left, right, 64, 63, **kw)
        want = O.compute_level(left, right, 64, 63, **kw)
    else:
        pl, pr = O.pyrdown(left), O.pyrdown(right)
        parent = O.compute_level(pl, pr, 32, 31, block=3, P1=20, P2=300, cost=1, post=1,
                                 want32=False)
        kw = dict(block=3, P1=20, P2=300, cost=1, post=3, paths=3, adaptive=1, refine=True,
                  min_conf=0.05)
        rc, d16, d32, cf = U.level(ws, left, right, 16, 63, parent.disp16, parent.conf, **kw)
        want = O.compute_level(left, right, 16, 63, parent.disp16, parent.conf, **kw)
    assert rc == 0 and want.rc == 0
    U.assert_same("disp16", d16, want.disp16)
    U.assert_same("disp32", d32, want.disp32)
    U.assert_same("conf", cf, want.conf)


@pytest.mark.parametrize("P2", [115, 116])
@pytest.mark.parametrize("kind", ["single", "coarse", "refine"])
def test_census_byte_volume_limit(ws, kind, P2):
    """Fixed P2 at the byte-volume limit: 24 + 2 * 115 < 255 runs on byte
    volumes with path costs up to 139; 24 + 2 * 116 does not and stays 16-bit.
    Both equal the oracle."""
    w, h = 176, 80
    left, right = pair(w, h, 60, seed=93 + P2)
    base = dict(P1=10, P2=P2, cost=1, adaptive=0)
    if kind == "single":
        kw = dict(block=5, post=3, paths=3, **base)
        rc, d16, d32, cf = U.single_scale(ws, left, right, 64, **kw)
        want = O.compute_single_scale(left, right, 64, **kw)
    elif kind == "coarse":
        kw = dict(block=3, post=1, paths=3, **base)
        rc, d16, d32, cf = U.level(ws, left, right, 64, 63, **kw)
        want = O.compute_level(left, right, 64, 63, **kw)
    else:
        pl, pr = O.pyrdown(left), O.pyrdown(right)
        parent = O.compute_level(pl, pr, 32, 31, block=3, post=1, want32=False, **base)
        kw = dict(block=3, post=3, paths=3, refine=True, min_conf=0.05, **base)
        rc, d16, d32, cf = U.level(ws, left, right, 16, 63, parent.disp16, parent.conf, **kw)
        want = O.compute_level(left, right, 16, 63, parent.disp16, parent.conf, **kw)
    assert rc == 0 and want.rc == 0
    U.assert_same("disp16", d16, want.disp16)
    U.assert_same("disp32", d32, want.disp32)
    U.assert_same("conf", cf, want.conf)
