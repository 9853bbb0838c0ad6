"""The batched form: N pairs per launch must equal N single-pair matches, bit for bit.

Every pair gets its own synthetic scene (seeded per pair).  The single-pair
results come from StereoMatcher (whose own parity with the oracle is tested in
test_gpu_parity.py); pair 0 is also compared with the oracle directly.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pyrsgm_amd import synth
from tests import gpu_util as U

pytestmark = pytest.mark.gpu

import torch

import pyrsgm_amd as P

PAD = 5  # extra columns, so every batched tensor has a row pitch != width


def batch_tensor(arrays, dtype, fill=0):
    n, (h, w) = len(arrays), arrays[0].shape if arrays else (0, 0)
    t = torch.full((n, h, w + PAD), fill, dtype=dtype, device="cuda")
    for i, a in enumerate(arrays):
        t[i, :, :w] = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t[:, :, :w]


def out_tensor(n, h, w, dtype, fill):
    return torch.full((n, h, w + PAD), fill, dtype=dtype, device="cuda")[:, :, :w]


def pairs(n, w, h, maxd, seed0):
    ps = [synth.make_pair(w, h, maxd, seed=seed0 + i) for i in range(n)]
    return [p.left for p in ps], [p.right for p in ps]


def singles(lefts, rights, params):
    return [U.matcher(l, r, params) for l, r in zip(lefts, rights)]


def run_batch(bm, L, R, n, h, w):
    d16 = out_tensor(n, h, w, torch.int16, 1234)
    d32 = out_tensor(n, h, w, torch.float32, -7.0)
    cf = out_tensor(n, h, w, torch.float32, -7.0)
    bm.compute(L, R, d16, d32, cf)
    torch.cuda.synchronize()
    return d16, d32, cf


CASES = [
    # (n, w, h, D, levels, cost, adaptive, median)
    (3, 200, 120, 64, 1, 1, 1, 1),
    (4, 241, 135, 128, 2, 1, 1, 1),
    (2, 160, 90, 64, 0, 0, 0, 1),
    (3, 192, 108, 64, 2, 0, 1, 0),
]


@pytest.mark.parametrize("n,w,h,D,levels,cost,adaptive,median", CASES)
def test_batch_equals_single_pairs(n, w, h, D, levels, cost, adaptive, median):
    lefts, rights = pairs(n, w, h, D - 1, seed0=500 + w)
    params = P.validate_params(P.Params(
        num_disparities=D, pyramid_levels=levels, cost_mode=P.CostMode(cost),
        p1=8 if cost else 300, p2=32 if cost else 2400, adaptive_p2=bool(adaptive),
        median_filter=bool(median)))
    bm = P.BatchMatcher(w, h, n, params)
    d16, d32, cf = run_batch(bm, batch_tensor(lefts, torch.uint8), batch_tensor(rights, torch.uint8),
                             n, h, w)
    want = singles(lefts, rights, params)
    for i in range(n):
        U.assert_same(f"pair {i} disp16", d16[i].cpu().numpy(), want[i][0])
        U.assert_same(f"pair {i} disp32", d32[i].cpu().numpy(), want[i][1])
        U.assert_same(f"pair {i} conf", cf[i].cpu().numpy(), want[i][2])
    ref = O.matcher_compute(lefts[0], rights[0], params)
    U.assert_same("pair 0 vs oracle", d16[0].cpu().numpy(), ref.disp16)
    bm.close()


def test_batch_views_with_different_strides():
    """Left and right batches with different pair strides (pyrDown per view)."""
    n, w, h, D = 3, 180, 100, 64
    lefts, rights = pairs(n, w, h, D - 1, seed0=77)
    params = P.validate_params(P.Params(num_disparities=D, pyramid_levels=2,
                                        cost_mode=P.CostMode.Census, p1=8, p2=32))
    L = batch_tensor(lefts, torch.uint8)
    Rbig = torch.zeros((n, h + 3, w + 11), dtype=torch.uint8, device="cuda")
    for i, r in enumerate(rights):
        Rbig[i, :h, :w] = torch.from_numpy(r).cuda()
    R = Rbig[:, :h, :w]
    bm = P.BatchMatcher(w, h, n, params)
    d16, d32, cf = run_batch(bm, L, R, n, h, w)
    want = singles(lefts, rights, params)
    for i in range(n):
        U.assert_same(f"pair {i} disp16", d16[i].cpu().numpy(), want[i][0])
        U.assert_same(f"pair {i} conf", cf[i].cpu().numpy(), want[i][2])
    bm.close()


def test_batch_graph_replays_match_eager():
    """Whole pyramid captured once; replayed several times on new inputs, with eager
    work and synchronisation in between; every replay equals the eager result."""
    n, w, h, D = 3, 200, 112, 128
    params = P.validate_params(P.Params(num_disparities=D, pyramid_levels=2,
                                        cost_mode=P.CostMode.Census, p1=8, p2=32,
                                        adaptive_p2=True))
    bm = P.BatchMatcher(w, h, n, params)
    eager = P.BatchMatcher(w, h, n, params)
    L = out_tensor(n, h, w, torch.uint8, 0)
    R = out_tensor(n, h, w, torch.uint8, 0)
    g16 = out_tensor(n, h, w, torch.int16, 1234)
    g32 = out_tensor(n, h, w, torch.float32, -7.0)
    gcf = out_tensor(n, h, w, torch.float32, -7.0)
    bm.capture(L, R, g16, g32, gcf)
    for rep in range(3):
        lefts, rights = pairs(n, w, h, D - 1, seed0=900 + 10 * rep)
        L.copy_(batch_tensor(lefts, torch.uint8))
        R.copy_(batch_tensor(rights, torch.uint8))
        torch.cuda.synchronize()
        bm.launch()
        torch.cuda.synchronize()
        e16, e32, ecf = run_batch(eager, L, R, n, h, w)   # eager work between replays
        U.assert_same(f"replay {rep} disp16", g16.cpu().numpy(), e16.cpu().numpy())
        U.assert_same(f"replay {rep} disp32", g32.cpu().numpy(), e32.cpu().numpy())
        U.assert_same(f"replay {rep} conf", gcf.cpu().numpy(), ecf.cpu().numpy())
        if rep == 0:
            ref = O.matcher_compute(lefts[1], rights[1], params)
            U.assert_same("replay 0 pair 1 vs oracle", g16[1].cpu().numpy(), ref.disp16)
    bm.close()
    eager.close()


def test_batch_partial_counts():
    """A matcher for 4 pairs run on its first 2, then all 4, then 1 (eager and
    captured): every pair equals the single-pair result, and pairs beyond the
    count are left untouched."""
    n, w, h, D = 4, 176, 104, 64
    lefts, rights = pairs(n, w, h, D - 1, seed0=4400)
    params = P.validate_params(P.Params(num_disparities=D, pyramid_levels=1,
                                        cost_mode=P.CostMode.Census, p1=8, p2=32,
                                        adaptive_p2=True))
    want = singles(lefts, rights, params)
    L = batch_tensor(lefts, torch.uint8)
    R = batch_tensor(rights, torch.uint8)
    bm = P.BatchMatcher(w, h, n, params)
    for count in (2, 4, 1):
        d16 = out_tensor(n, h, w, torch.int16, 1234)
        bm.compute(L, R, d16, count=count)
        torch.cuda.synchronize()
        for i in range(n):
            if i < count:
                U.assert_same(f"count {count} pair {i}", d16[i].cpu().numpy(), want[i][0])
            else:
                assert bool((d16[i] == 1234).all()), f"count {count} wrote pair {i}"
    g16 = out_tensor(n, h, w, torch.int16, 1234)
    bm.capture(L, R, g16, count=3)
    for rep in range(2):
        g16.fill_(1234)
        torch.cuda.synchronize()
        bm.launch()
        torch.cuda.synchronize()
        for i in range(3):
            U.assert_same(f"graph rep {rep} pair {i}", g16[i].cpu().numpy(), want[i][0])
        assert bool((g16[3] == 1234).all())
    bm.close()


@pytest.mark.parametrize("n,w,h,D,levels", [
    (16, 480, 272, 64, 1),   # refinement level 480x272x16: 4,096 four-line waves -> paired lines
    (32, 480, 272, 32, 1),   # coarse level 240x136x16 with 32 pairs -> paired lines
])
def test_batch_paired_lines_equal_single_pairs(n, w, h, D, levels):
    """Launches with many lines run two lines per lane (16-bit halves); single
    pairs never do.  Every pair must still equal its single-pair result."""
    lefts, rights = pairs(n, w, h, D - 1, seed0=7000 + D)
    params = P.validate_params(P.Params(num_disparities=D, pyramid_levels=levels,
                                        cost_mode=P.CostMode.Census, p1=8, p2=32,
                                        adaptive_p2=True))
    L = batch_tensor(lefts, torch.uint8)
    R = batch_tensor(rights, torch.uint8)
    bm = P.BatchMatcher(w, h, n, params)
    d16, d32, cf = run_batch(bm, L, R, n, h, w)
    bm.close()
    for i in (0, n // 2, n - 1):
        want = U.matcher(lefts[i], rights[i], params)
        U.assert_same(f"pair {i} disp16", d16[i].cpu().numpy(), want[0])
        U.assert_same(f"pair {i} conf", cf[i].cpu().numpy(), want[2])
    ref = O.matcher_compute(lefts[1], rights[1], params)
    U.assert_same("pair 1 vs oracle", d16[1].cpu().numpy(), ref.disp16)


@pytest.mark.parametrize("census_rows,paired,xcd_remap,side_stream,fused,modular",
                         [(2, 1, 0, 0, 1, -1), (8, 0, 1, 1, 1, 1), (4, -1, -1, 1, -1, 0),
                          (2, -1, 1, 0, 0, 1), (4, 1, -1, 1, 0, 0), (8, -1, 0, 1, 1, -1),
                          (4, 0, -1, 1, 0, 0), (4, 0, -1, 1, 1, 0)])
def test_batch_tunings_equal_single_pairs(census_rows, paired, xcd_remap, side_stream, fused,
                                          modular):
    """Every launch shape (census rows per block, paired lines forced on or off,
    XCD remap, side stream, winner-take-all fused into the R->L lines or not,
    shifting or modular refinement step) gives the same bits as the pairs
    matched alone."""
    n, w, h, D = 3, 241, 135, 128
    lefts, rights = pairs(n, w, h, D - 1, seed0=900)
    params = P.validate_params(P.Params(num_disparities=D, pyramid_levels=2,
                                        cost_mode=P.CostMode.Census, p1=8, p2=32,
                                        adaptive_p2=True, median_filter=True))
    bm = P.BatchMatcher(w, h, n, params)
    bm.set_tuning(census_rows, paired, xcd_remap, side_stream, fused, modular)
    assert bm.tuning() == {"census_rows": census_rows, "paired": paired, "xcd_remap": xcd_remap,
                           "side_stream": side_stream, "fused": fused, "modular": modular}
    print("case: batch tuning", bm.tuning(), flush=True)
    d16, d32, cf = run_batch(bm, batch_tensor(lefts, torch.uint8), batch_tensor(rights, torch.uint8),
                             n, h, w)
    want = singles(lefts, rights, params)
    for i in range(n):
        U.assert_same(f"pair {i} disp16", d16[i].cpu().numpy(), want[i][0])
        U.assert_same(f"pair {i} conf", cf[i].cpu().numpy(), want[i][2])
    bm.close()


FUSED_CASES = [
    # (n, w, h, D, levels, adaptive, num_paths, paired): the fused winner-take-all
    # at refinement levels (modular, paired and plain R->L lines) and at a
    # 16-label coarse level (D=64 over two halvings), fixed and adaptive P2,
    # with fewer paths (no T->B / no L->R volume read).
    (3, 241, 135, 128, 2, 1, 3, 0),
    (3, 241, 135, 128, 2, 1, 3, 1),
    (2, 203, 117, 64, 2, 0, 3, 0),
    (2, 203, 117, 64, 2, 1, 3, 1),
    (2, 160, 96, 64, 2, 1, 2, 0),
    (2, 160, 96, 64, 1, 0, 1, 1),
]


@pytest.mark.parametrize("n,w,h,D,levels,adaptive,num_paths,paired", FUSED_CASES)
def test_fused_wta_equals_single_pairs(n, w, h, D, levels, adaptive, num_paths, paired):
    """Winner-take-all finished by the R->L lines (forced on) gives the same bits
    as separate winner-take-all launches (the pairs matched alone); pair 0 is
    also compared with the oracle."""
    lefts, rights = pairs(n, w, h, D - 1, seed0=1700 + w + levels)
    params = P.validate_params(P.Params(num_disparities=D, pyramid_levels=levels,
                                        cost_mode=P.CostMode.Census, p1=8, p2=32,
                                        adaptive_p2=bool(adaptive), median_filter=True,
                                        num_paths=num_paths))
    bm = P.BatchMatcher(w, h, n, params)
    bm.set_tuning(4, paired, -1, 1, 1)
    print(f"case: fused wta {n}x{w}x{h} D={D} levels={levels} adaptive={adaptive} "
          f"paths={num_paths} paired={paired}", flush=True)
    d16, d32, cf = run_batch(bm, batch_tensor(lefts, torch.uint8), batch_tensor(rights, torch.uint8),
                             n, h, w)
    want = singles(lefts, rights, params)
    for i in range(n):
        U.assert_same(f"pair {i} disp16", d16[i].cpu().numpy(), want[i][0])
        U.assert_same(f"pair {i} disp32", d32[i].cpu().numpy(), want[i][1])
        U.assert_same(f"pair {i} conf", cf[i].cpu().numpy(), want[i][2])
    ref = O.matcher_compute(lefts[0], rights[0], params)
    assert ref.rc == 0
    U.assert_same("pair 0 disp16 vs oracle", d16[0].cpu().numpy(), ref.disp16)
    U.assert_same("pair 0 conf vs oracle", cf[0].cpu().numpy(), ref.conf)
    bm.close()


def test_batch_tune_keeps_results():
    n, w, h, D = 4, 200, 120, 64
    lefts, rights = pairs(n, w, h, D - 1, seed0=950)
    params = P.validate_params(P.Params(num_disparities=D, pyramid_levels=1,
                                        cost_mode=P.CostMode.Census, p1=8, p2=32,
                                        adaptive_p2=True))
    bm = P.BatchMatcher(w, h, n, params)
    L, R = batch_tensor(lefts, torch.uint8), batch_tensor(rights, torch.uint8)
    d16 = out_tensor(n, h, w, torch.int16, 0)
    print("case: batch tune", flush=True)
    best, ms = bm.tune(L, R, d16, reps=2)
    assert ms > 0 and bm.tuning() == best and best["census_rows"] in (2, 4, 8)
    with pytest.raises(ValueError):
        bm.set_tuning(census_rows=3)
    with pytest.raises(ValueError):
        bm.set_tuning(modular=2)
    bm.set_tuning(**best)
    d16, d32, cf = run_batch(bm, L, R, n, h, w)
    want = singles(lefts, rights, params)
    for i in range(n):
        U.assert_same(f"tuned pair {i} disp16", d16[i].cpu().numpy(), want[i][0])
    # A new matcher of the same key starts with the tuned shapes.
    bm2 = P.BatchMatcher(w, h, n, params)
    assert bm2.tuning() == best
    bm2.close()
    bm.close()
