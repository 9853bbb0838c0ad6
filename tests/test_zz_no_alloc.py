"""pyrsgm_compute_single_scale_noalloc against pyrsgm_compute_single_scale.

Two calls in a row through the C-ABI, the second on a smaller input (the
scratch array keeps the first call's contents), each compared bit for bit
with the workspace entry point on the same input; then a size whose workspace
exceeds the fixed array is rejected with -1.
"""
import pytest
import torch

import pyrsgm_amd as P
from pyrsgm_amd import synth
from tests import gpu_util as U

pytestmark = pytest.mark.gpu

CASES = [
    # (w, h, D, minD, cost, post, adaptive, P1, P2): the 640x480 D=64 BT pair, then smaller
    (640, 480, 64, 0, 0, 3, 0, 200, 800),
    (200, 120, 32, -3, 1, 1, 1, 8, 32),
]


def run_noalloc(left, right, D, minD, cost, post, adaptive, P1, P2):
    h, w = left.shape
    L, R = U.Img.upload(left), U.Img.upload(right)
    d16 = U.Img(h, w, torch.int16, 1234)
    d32 = U.Img(h, w, torch.float32, -7.0)
    cf = U.Img(h, w, torch.float32, -7.0)
    rc = P.compute_single_scale_noalloc(
        L.ptr, R.ptr, d16.ptr, d32.ptr, cf.ptr, L.pitch, R.pitch, d16.pitch, d32.pitch, cf.pitch,
        w, h, D, minD, 5, P1, P2, 10, 1, 63, cost, post, 3, adaptive, 3.041, 16.137, 0.601, 1.77,
        U.stream())
    U.sync()
    return rc, d16.numpy(), d32.numpy(), cf.numpy()


def test_noalloc_matches_workspace_entry_point():
    ws = P.create_workspace()
    try:
        for (w, h, D, minD, cost, post, adaptive, P1, P2) in CASES:
            pair = synth.make_pair(w, h, D - 4, seed=w + D)
            print(f"case: noalloc {w}x{h} D={D} minD={minD} cost={cost}", flush=True)
            rc, a16, a32, ac = run_noalloc(pair.left, pair.right, D, minD, cost, post, adaptive,
                                           P1, P2)
            assert rc == 0
            rc2, b16, b32, bc = U.single_scale(ws, pair.left, pair.right, D, minD=minD, block=5,
                                               P1=P1, P2=P2, uniq=10, disp12=1, cap=63, cost=cost,
                                               post=post, paths=3, adaptive=adaptive)
            assert rc2 == 0
            U.assert_same("disp16", a16, b16)
            U.assert_same("disp32", a32, b32)
            U.assert_same("conf", ac, bc)
            assert (a16 != (minD - 1) * 16).any()
    finally:
        P.destroy_workspace(ws)


def test_noalloc_rejects_sizes_beyond_the_scratch_array():
    # 1280x720 with 256 disparities needs far more than the fixed array.
    w, h, D = 1280, 720, 256
    L = U.Img(h, w, torch.uint8)
    d16 = U.Img(h, w, torch.int16, 1234)
    rc = P.compute_single_scale_noalloc(
        L.ptr, L.ptr, d16.ptr, None, None, L.pitch, L.pitch, d16.pitch, 0, 0, w, h, D, 0, 5,
        8, 32, 10, 1, 63, 1, 3, 3, 0, 3.041, 16.137, 0.601, 1.77, U.stream())
    U.sync()
    assert rc == -1
    assert (d16.numpy() == 1234).all()   # nothing was launched
