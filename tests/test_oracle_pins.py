"""Pins for the CPU oracle, stage by stage, against checks that do not reuse its code.

Each test states what pins the stage: the reference's own known-answer test
(pyrDown), an independent algorithm (cumulative sums for the window sums, the
textbook O(D^2) recurrence, numpy argmin, sorting for the median), closed forms
(census of a ramp, BT of identical rows, parabola vertices), or the
reference's acceptance thresholds (tests/matcher_test.cpp, tests/pyramid_smoke.cpp).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

MAX = 32767
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rng(seed):
    return np.random.default_rng(seed)


# ------------------------------------------------------------------ pyrDown
def pyrdown_independent(src):
    """Separable [1 4 6 4 1] correlation with reflect-101 (numpy 'reflect' mode),
    even samples, (sum + 128) >> 8 -- the KAT of tests/matcher_test.cpp:74-101."""
    k = np.array([1, 4, 6, 4, 1], np.int64)
    p = np.pad(src.astype(np.int64), 2, mode="reflect")
    h, w = src.shape
    rows = sum(k[i] * p[i:i + h, :] for i in range(5))
    full = sum(k[i] * rows[:, i:i + w] for i in range(5))
    return ((full[::2, ::2] + 128) >> 8).astype(np.uint8)


def kat_input(w, h):
    xs, ys = np.meshgrid(np.arange(w), np.arange(h))
    return ((xs * 13 + ys * 29 + (xs * ys) // 7) & 0xFF).astype(np.uint8)


@pytest.mark.parametrize("w,h", [(64, 48), (64, 33), (97, 48), (97, 33)])
def test_pyrdown_reference_kat(w, h):
    src = kat_input(w, h)
    got = O.pyrdown(src)
    np.testing.assert_array_equal(got, pyrdown_independent(src))
    fx = np.load(os.path.join(GOLDEN, "pyrdown_kat.npz"))
    np.testing.assert_array_equal(got, fx[f"out_{w}x{h}"])


def test_pyrdown_constant_and_size_checks():
    src = np.full((9, 7), 77, np.uint8)
    assert (O.pyrdown(src) == 77).all()
    with pytest.raises(ValueError):
        O.pyrdown(np.zeros((1, 5), np.uint8))


# ------------------------------------------------------------------- census
def census_independent(img):
    h, w = img.shape
    p = np.pad(img.astype(np.int32), 2, mode="edge")
    c = img.astype(np.int32)
    out = np.zeros((h, w), np.uint64)
    for dy in range(5):
        for dx in range(5):
            bit = (p[dy:dy + h, dx:dx + w] < c).astype(np.uint64)
            out = (out << np.uint64(1)) | bit
    return out.astype(np.uint32)


def test_census_matches_independent_transform():
    img = rng(1).integers(0, 256, (23, 37), dtype=np.uint8)
    np.testing.assert_array_equal(O.census(img), census_independent(img))


def test_census_closed_forms():
    # Constant image: no neighbour is darker.
    assert (O.census(np.full((8, 8), 9, np.uint8)) == 0).all()
    # Horizontal ramp I = x: in every window row only dx = -2, -1 are darker,
    # i.e. bits 11000 per row (interior pixels).
    ramp = np.tile(np.arange(20, dtype=np.uint8), (10, 1))
    expected = int("11000" * 5, 2)
    assert (O.census(ramp)[:, 2:-2] == expected).all()


def test_census_cost_range_and_borders():
    r = rng(2)
    l = r.integers(0, 256, (12, 40), dtype=np.uint8)
    cl, cr = O.census(l), O.census(l)
    vol = O.census_cost(cl, cr, 16)
    assert vol.min() >= 0 and vol.max() <= 24
    assert (vol[:, :, 0] == 0).all()                    # identical views at d = 0
    assert (vol[:, 3, 4:] == 24).all()                  # x - d < 0 -> 24
    hd = np.array([bin(int(a ^ b)).count("1") for a, b in zip(cl[5, 10:20], cr[5, 7:17])])
    np.testing.assert_array_equal(vol[5, 10:20, 3], hd)


# ----------------------------------------------------------------------- BT
def bt_interval(v, l, r):
    hl, hr = (v + l) // 2, (v + r) // 2
    return min(hl, hr, v), max(hl, hr, v)


def bt_brute(u, ul, ur, v, vl, vr):
    """Birchfield-Tomasi by enumeration: distance from a sample to the other
    side's half-pixel interval, minimised over every integer in it, both ways."""
    u0, u1 = bt_interval(u, ul, ur)
    v0, v1 = bt_interval(v, vl, vr)
    a = min(abs(u - t) for t in range(v0, v1 + 1))
    b = min(abs(v - t) for t in range(u0, u1 + 1))
    return min(a, b)


def test_bt_cost_by_enumeration():
    r = rng(3)
    h, w, D = 5, 24, 8
    left = r.integers(0, 256, (h, w), dtype=np.uint8)
    right = r.integers(0, 256, (h, w), dtype=np.uint8)
    vol = O.bt_cost(left, right, D)
    g1, g2 = O.bt_grad(left).astype(int), O.bt_grad(right).astype(int)
    L, R = left.astype(int), right.astype(int)
    for y in range(h):
        for x in range(w):
            for d in range(D):
                x2 = x - d
                if x2 < 0:
                    assert vol[y, x, d] == 255
                    continue
                xm, xp = max(x - 1, 0), min(x + 1, w - 1)
                x2m, x2p = max(x2 - 1, 0), min(x2 + 1, w - 1)
                cg = bt_brute(g1[y, x], g1[y, xm], g1[y, xp], g2[y, x2], g2[y, x2m], g2[y, x2p])
                ci = bt_brute(L[y, x], L[y, xm], L[y, xp], R[y, x2], R[y, x2m], R[y, x2p])
                assert vol[y, x, d] == min(255, cg + (ci >> 2)), (y, x, d)


def test_bt_properties():
    r = rng(4)
    img = r.integers(0, 256, (6, 30), dtype=np.uint8)
    vol = O.bt_cost(img, img, 4)
    assert (vol[:, :, 0] == 0).all()                     # identical views
    # BT never exceeds the absolute difference of the two samples.
    g = O.bt_grad(img).astype(int)
    ad = np.abs(g[:, 5:] - g[:, 3:-2]) + (np.abs(img[:, 5:].astype(int) - img[:, 3:-2]) >> 2)
    assert (vol[:, 5:, 2] <= np.minimum(255, ad)).all()
    # Gradient channel: clipped to [0, 2*ftzero]; flat image -> ftzero.
    assert O.bt_grad(np.full((4, 4), 50, np.uint8)).max() == 63
    assert O.bt_grad(img, 15).max() <= 30


def test_box_sums_against_cumulative_sums():
    r = rng(5)
    vol = r.integers(0, 256, (9, 13, 4)).astype(np.int16)
    for block in (1, 3, 5):
        half = block // 2
        p = np.pad(vol.astype(np.int64), ((half, half), (half, half), (0, 0)), mode="edge")
        cs = np.cumsum(np.cumsum(np.pad(p, ((1, 0), (1, 0), (0, 0))), 0), 1)
        h, w = vol.shape[:2]
        s = cs[block:block + h, block:block + w] - cs[:h, block:block + w] \
            - cs[block:block + h, :w] + cs[:h, :w]
        np.testing.assert_array_equal(O.box(vol, block), np.minimum(s, MAX))


def test_box_saturates_each_pass():
    vol = np.full((3, 4, 2), 30000, np.int16)
    out = O.box(vol, 3)
    assert (out == MAX).all()


# ------------------------------------------------------------ adaptive P2
def test_p2_table_closed_forms():
    lut, on = O.p2_lut(800, 1, 3.041, 16.137, 0.601, 1.77)
    assert on
    assert lut[0] == int(800 * 1.77 + 0.5)            # (1+alpha) clamped by upper
    assert lut[255] == int(800 * 0.601 + 0.5)         # exp -> 0: lower
    assert (np.diff(lut) <= 0).all()
    lut2, on2 = O.p2_lut(800, 1, 0.0, 16.0, 1.0, 1.0)  # alpha 0 and lower >= 1: constant
    assert not on2 and (lut2 == 800).all()
    lut3, _ = O.p2_lut(32, 1, 2.0, 14.0, 0.75, 2.0)
    e = np.exp(-np.arange(256, dtype=np.float64) / 14.0)
    ref = np.clip(32 * (0.75 + 2.25 * e), 32 * 0.75, 64)
    assert np.abs(lut3 - np.floor(ref + 0.5)).max() <= 1


# ------------------------------------------------------------- recurrences
def textbook_path(C, P1, P2, lut=None, img=None):
    """Hirschmueller's recurrence along axis 1 with every transition priced
    explicitly: L(x,d) = C + min_k(L(x-1,k) + pen(|d-k|)) - min_k L(x-1,k)."""
    h, w, D = C.shape
    L = np.zeros((h, w, D), np.int64)
    L[:, 0] = C[:, 0]
    dd = np.abs(np.arange(D)[:, None] - np.arange(D)[None, :])
    for x in range(1, w):
        prev = L[:, x - 1]
        p2 = np.full(h, P2) if lut is None else lut[np.abs(img[:, x].astype(int) - img[:, x - 1])]
        pen = np.where(dd == 0, 0, np.where(dd == 1, P1, p2[:, None, None]))
        best = (prev[:, None, :] + pen).min(axis=2)
        L[:, x] = C[:, x] + best - prev.min(axis=1)[:, None]
    return L


@pytest.mark.parametrize("P1,P2,adaptive", [(3, 11, False), (5, 5, False), (2, 40, True)])
def test_single_scale_lr_matches_textbook(P1, P2, adaptive):
    r = rng(6)
    C = r.integers(0, 60, (7, 30, 16)).astype(np.int16)
    img = r.integers(0, 256, (7, 30), dtype=np.uint8)
    lut = O.p2_lut(P2, 1, 2.0, 10.0, 0.6, 1.8)[0] if adaptive else None
    got = O.path_lr(C, img, P1, P2, lut=lut)
    np.testing.assert_array_equal(got, textbook_path(C, P1, P2, lut, img))


def test_recurrence_special_cases():
    r = rng(7)
    C = r.integers(0, 50, (4, 20, 16)).astype(np.int16)
    img = np.zeros((4, 20), np.uint8)
    np.testing.assert_array_equal(O.path_lr(C, img, 0, 0), C)      # no smoothness: L = C
    tb = O.path_tb(C, img, 0, 0)
    np.testing.assert_array_equal(tb, C)
    L = O.path_lr(C, img, 4, 20)
    assert (L >= C).all() and (L <= C.astype(int) + 20).all()        # C <= L <= C + P2


def test_single_scale_tb_is_lr_of_transpose():
    r = rng(8)
    C = r.integers(0, 60, (9, 11, 16)).astype(np.int16)
    img = r.integers(0, 256, (9, 11), dtype=np.uint8)
    tb = O.path_tb(C, img, 3, 17)
    lrT = O.path_lr(np.ascontiguousarray(C.transpose(1, 0, 2)), np.ascontiguousarray(img.T), 3, 17)
    np.testing.assert_array_equal(tb, lrT.transpose(1, 0, 2))


def textbook_pyramid_lr(C, base, valid, P1, P2, minD, maxD):
    """Pyramid L->R in real-disparity space: every label is a real disparity
    base+d; a transition from the previous pixel prices |real - real'| only
    over the previous pixel's labels that carry a path (cost < MAX)."""
    h, w, D = C.shape
    out = np.zeros_like(C, dtype=np.int64)
    for y in range(h):
        prev = None  # dict real -> L
        for x in range(w):
            reals = {}
            for d in range(D):
                real = int(base[y, x]) + d
                ok = (valid[y, x] and minD <= real <= maxD and 0 <= x - real < w
                      and C[y, x, d] < MAX)
                reals[d] = (real, ok)
            if not any(ok for _, ok in reals.values()):
                out[y, x] = MAX
                prev = None
                continue
            L = np.full(D, MAX, np.int64)
            if prev is None:
                for d, (real, ok) in reals.items():
                    if ok:
                        L[d] = C[y, x, d]
            else:
                mprev = min(prev.values())
                for d, (real, ok) in reals.items():
                    if not ok:
                        continue
                    best = MAX
                    for rp, lv in prev.items():
                        if lv >= MAX:
                            continue
                        dr = abs(rp - real)
                        best = min(best, lv + (0 if dr == 0 else P1 if dr == 1 else P2))
                    best = min(best, mprev + P2, MAX)
                    L[d] = MAX if best >= MAX else min(MAX, C[y, x, d] + best - mprev)
            out[y, x] = L
            cur = {reals[d][0]: int(L[d]) for d in range(D)}
            prev = cur if L.min() < MAX else None
    return out


def test_pyramid_lr_matches_real_disparity_textbook():
    r = rng(9)
    h, w, D = 4, 40, 16
    C = r.integers(0, 25, (h, w, D)).astype(np.int16)
    base = np.clip(r.integers(-3, 4, (h, w)).cumsum(axis=1) + 5, 0, 20).astype(np.int16)
    valid = (r.random((h, w)) > 0.1).astype(np.uint8)
    C[r.random((h, w, D)) < 0.05] = MAX
    maxD = 30
    got = O.path_pyr("lr", C, base, valid, np.zeros((h, w), np.uint8), 3, 12, maxD)
    np.testing.assert_array_equal(got, textbook_pyramid_lr(C, base, valid, 3, 12, 0, maxD))


# --------------------------------------------------------- WTA and post
def test_wta_uniqueness_subpixel_against_numpy():
    """With P1 = P2 = 0 the R->L path equals the cost, so the summed cost is
    S = min(MAX, acc + C) and WTA can be recomputed independently."""
    r = rng(10)
    h, w, D = 6, 70, 32
    C = r.integers(0, 40, (h, w, D)).astype(np.int16)
    acc = r.integers(0, 200, (h, w, D)).astype(np.int16)
    img = np.zeros((h, w), np.uint8)
    out = O.path_rl_wta(C, img, 0, 0, uniq=15, acc=acc)
    S = np.minimum(MAX, acc.astype(np.int64) + C)
    best = S.argmin(axis=2)                      # first minimum = smallest d
    mn = S.min(axis=2)
    for y in range(h):
        for x in range(w - 1):
            if x < D:
                assert out.disp[y, x] == -16
                continue
            b, m = best[y, x], mn[y, x]
            alts = [S[y, x, d] for d in range(D) if abs(d - b) > 1]
            alt = min(alts)
            unique = not any(S[y, x, d] * 85 < m * 100 for d in range(D) if abs(d - b) > 1)
            if not unique:
                assert out.disp[y, x] == -16
                continue
            if 0 < b < D - 1:
                c0, c1, c2 = (int(v) for v in S[y, x, b - 1:b + 2])
                den = max(c0 + c2 - 2 * c1, 1)
                num = (c0 - c2) * 16 + den
                q = abs(num) // (2 * den) * (1 if num >= 0 else -1)   # C truncation
                want = b * 16 + q
            else:
                want = b * 16
            assert out.disp[y, x] == want, (y, x)
            cf = np.float32(np.float32(alt) - np.float32(m)) / np.float32(max(m, 1))
            assert out.conf[y, x] == np.clip(cf, 0, 1)
    assert (out.disp[:, w - 1] == -16).all()


def test_subpixel_recovers_parabola_vertex():
    """For costs sampled from a parabola with vertex at best+delta, the
    reference formula yields 16*best + trunc(16*delta + 1/2)."""
    for k in range(-7, 8):
        delta = k / 16.0
        a = 64.0
        c = [a * (dd - delta) ** 2 + 1000 for dd in (-1, 0, 1)]
        c0, c1, c2 = (int(round(v)) for v in c)
        if not (c1 <= c0 and c1 <= c2):
            continue
        den = max(c0 + c2 - 2 * c1, 1)
        num = (c0 - c2) * 16 + den
        q = int(num / (2 * den))
        assert q == int(16 * delta + 0.5)


def test_disp2_order_free_key_equals_sequential_update():
    r = rng(11)
    h, w, D = 5, 80, 16
    C = r.integers(0, 30, (h, w, D)).astype(np.int16)
    out = O.path_rl_wta(C, np.zeros((h, w), np.uint8), 2, 9, uniq=5)
    # Recompute every accepted pixel's (min cost, best d) from the summed costs.
    S = out.acc.astype(np.int64)
    keys = np.full((h, w), 0xFFFFFFFF, np.uint64)
    for y in range(h):
        for x in range(w - 1):
            if out.disp[y, x] == -16:
                continue
            b = int(S[y, x].argmin())
            m = int(S[y, x].min())
            x2 = x - b
            if 0 <= x2 < w:
                k = ((m + 32768) << 16) | (0xFFFF - x)
                keys[y, x2] = min(keys[y, x2], k)
    for y in range(h):
        for x2 in range(w):
            k = int(keys[y, x2])
            if k == 0xFFFFFFFF:
                assert out.disp2[y, x2] == -16 and out.disp2_cost[y, x2] == MAX
            else:
                xw = 0xFFFF - (k & 0xFFFF)
                assert out.disp2[y, x2] == xw - x2
                assert out.disp2_cost[y, x2] == (k >> 16) - 32768


def test_lr_checks_hand_cases():
    disp = np.full((1, 20), -16, np.int16)
    disp2 = np.full((1, 20), -16, np.int16)
    conf = np.ones((1, 20), np.float32)
    disp[0, 10] = 5 * 16          # d0 = d1 = 5 -> right pixel 5
    disp[0, 12] = 5 * 16 + 8      # d0 = 5, d1 = 6 -> right pixels 7, 6
    disp[0, 15] = 3 * 16          # right pixel 12 missing (invalid marker)
    disp2[0, 5] = 7               # |7 - 5| = 2 > 1
    disp2[0, 7] = 5               # consistent with d0 = 5
    disp2[0, 6] = 9               # inconsistent with d1 = 6
    d, c = O.lr_check(disp, disp2, conf)                 # pseudo: needs both bad
    assert d[0, 10] == -16 and c[0, 10] == 0
    assert d[0, 12] == 88 and d[0, 15] == 48
    d, c = O.lr_check(disp, disp2, conf, strict=True)    # strict: needs one good
    assert d[0, 10] == -16 and d[0, 12] == 88 and d[0, 15] == -16 and c[0, 15] == 0


def test_median_against_sort():
    r = rng(12)
    d = r.integers(-16, 300, (7, 9)).astype(np.int16)
    got = O.median3x3(d)
    h, w = d.shape
    for y in range(h):
        for x in range(w):
            win = d[max(0, y - 1):y + 2, max(0, x - 1):x + 2].ravel()
            assert got[y, x] == np.sort(win)[len(win) // 2]


# ------------------------------------------------------- pyramid helpers
def test_refine_labels_hand_cases():
    pd = np.array([[80, 88, 84, -16, 80]], np.int16)
    pc = np.array([[1, 1, 1, 1, 0.01]], np.float32)
    base, valid = O.prepare_refine_labels(pd, pc, 10, 1, maxD=255, min_conf=0.05)
    # parent 5.0 px -> centre 10, residual 0 -> base 10 - 8 + 1; but x < 15 clamps to 0.
    assert (base[0, :] == 0).all()
    pd2 = np.full((1, 60), 80, np.int16)
    pd2[0, 50] = 88
    pd2[0, 51] = 84
    base, valid = O.prepare_refine_labels(pd2, np.ones((1, 60), np.float32), 120, 1, maxD=255)
    assert base[0, 60] == 3 and valid[0, 60] == 1      # 2*80=160: centre 10, res 0 -> 3
    assert base[0, 100] == 4                            # 2*88=176: centre 11, res 0 -> 4
    assert base[0, 102] == 3                            # 2*84=168: centre 11, res -8 -> 3
    base, valid = O.prepare_refine_labels(pd, pc, 10, 1, maxD=255)
    assert valid[0, 6] == 0 and valid[0, 8] == 0        # invalid parent, low confidence


def test_level_geometry_reference_case():
    """tests/matcher_test.cpp:266-281: 961x541, D=128, 3 extra levels."""
    import pyrsgm_amd as P
    p = P.validate_params(P.Params(num_disparities=128, pyramid_levels=3))
    info = O.level_geometry(961, 541, p)
    assert len(info) == 4
    assert (info[3]["width"], info[3]["height"]) == (121, 68)
    assert info[3]["refine"] == 0 and info[3]["num_labels"] == 16 and info[3]["max_disparity"] == 15
    assert info[0]["refine"] == 1 and info[0]["num_labels"] == 16
    assert info[2]["block_size"] == 3 and info[1]["block_size"] == 5


# ----------------------------------- end to end: the reference's thresholds
def pattern(w, h, d, right):
    xs = np.arange(w)[None, :] + (d if right else 0)
    ys = np.arange(h)[:, None]
    return ((xs * 37 + ys * 17 + ((xs // 7) ^ (ys // 5)) * 29) & 0xFF).astype(np.uint8)


def sample_texture(xf, y):
    def tex(sx, sy):
        return (sx * 37 + sy * 17 + ((sx // 7) ^ (sy // 5)) * 29) & 0xFF
    x0 = np.floor(xf).astype(np.int64)
    f = (xf - x0).astype(np.float32)
    v = (np.float32(1) - f) * tex(x0, y).astype(np.float32) + f * tex(x0 + 1, y).astype(np.float32)
    return (v + np.float32(0.5)).astype(np.uint8)


def disp_stats(d32, gt, margin):
    h, w = d32.shape
    sub = d32[margin:h - margin, margin:w - margin]
    g = gt[margin:h - margin, margin:w - margin]
    valid = sub > 0
    err = np.abs(sub - g)[valid]
    return valid.mean(), (err <= 1).mean() if err.size else 0.0, err.mean() if err.size else 0.0


@pytest.mark.parametrize("levels", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("census", [False, True])
def test_reference_acceptance_thresholds(levels, census):
    """tests/matcher_test.cpp:198-247 run on the oracle."""
    import pyrsgm_amd as P
    w, h = 640, 480
    p = P.validate_params(P.Params(num_disparities=64, pyramid_levels=levels,
                                   cost_mode=P.CostMode(int(census)),
                                   p1=8 if census else 300, p2=32 if census else 2400))
    left, right = pattern(w, h, 40, False), pattern(w, h, 40, True)
    out = O.matcher_compute(left, right, p)
    dens, good, _ = disp_stats(out.disp32, np.full((h, w), 40.0, np.float32), 64 + 8)
    assert dens > 0.5 and good > 0.95
    inv = out.disp16 == -16
    assert inv.any() and (out.disp32[inv] == 0).all()
    ok = ~inv & (out.disp32 != 0)
    assert np.abs(out.disp16[ok] / 16.0 - out.disp32[ok]).max() <= 1e-6
    ys = np.arange(h)[:, None] * np.ones((1, w), np.int64)
    xs = np.arange(w, dtype=np.float32)[None, :] * np.ones((h, 1), np.float32)
    sl = sample_texture(xs, ys)
    sr = sample_texture((xs + np.float32(12.0)) / np.float32(1.0 - 0.06), ys)
    gt = np.float32(12.0) + np.float32(0.06) * xs
    out = O.matcher_compute(sl, sr, p)
    _, good, mae = disp_stats(out.disp32, gt, 64 + 8)
    assert good > 0.9 and mae < 0.5


def test_reference_pyramid_smoke_chain():
    """tests/pyramid_smoke.cpp:93-113: a real coarse result refined over four levels."""
    sizes = [(40, 16, 2), (80, 32, 4), (160, 64, 8), (320, 128, 16), (640, 256, 32)]
    maxd = [31, 63, 127, 255, 511]
    kw = dict(block=3, P1=24, P2=96, uniq=0, disp12=1, cap=63, cost=O.COST_BT, post=1, paths=3,
              min_conf=0.0, adaptive=1, alpha=2.0, beta=14.0, lower=0.75, upper=2.0)
    parent = None
    for i, (w, h, d) in enumerate(sizes):
        l, r = pattern(w, h, d, False), pattern(w, h, d, True)
        if i == 0:
            out = O.compute_level(l, r, 32, maxd[i], refine=False, **kw)
        else:
            out = O.compute_level(l, r, 16, maxd[i], parent.disp16, parent.conf, refine=True, **kw)
        assert out.rc == 0
        v = out.disp32[out.disp32 > 0]
        tol = 2.0 if i == 4 else 1.5
        assert v.size > 0 and (np.abs(v - d) <= tol).sum() >= v.size / 2
        parent = out


# ------------------------------------- BT gradient: closed form + ramp
def bt_grad_numpy(img, cap):
    """sgm_kernels.cu:94-118 as array arithmetic over an edge-padded image:
    2*(row[x+1]-row[x-1]) + the same difference one row up and one row down,
    clipped to +-ftzero (ftzero = max(cap, 15) | 1), offset by ftzero and
    stored as a byte (so it wraps above 255)."""
    ft = max(cap, 15) | 1
    p = np.pad(img.astype(np.int64), 1, mode="edge")
    dx = p[:, 2:] - p[:, :-2]                    # [h + 2, w]
    g = 2 * dx[1:-1] + dx[:-2] + dx[2:]
    return ((np.clip(g, -ft, ft) + ft) & 0xFF).astype(np.uint8)


@pytest.mark.parametrize("cap", [1, 15, 31, 63, 127, 128, 200])
def test_bt_grad_against_numpy(cap):
    img = rng(20 + cap).integers(0, 256, (19, 33), dtype=np.uint8)
    np.testing.assert_array_equal(O.bt_grad(img, cap), bt_grad_numpy(img, cap))


def test_bt_grad_ramp_closed_form():
    """Interior of I = a*x: every horizontal difference is 2a, so the
    gradient is 2*2a + 2a + 2a = 8a, offset by ftzero = 63 (no clipping for
    a <= 7); the left and right border columns see one-sided differences."""
    for a in range(0, 8):
        img = np.tile((np.arange(30) * a).astype(np.uint8), (6, 1))
        g = O.bt_grad(img, 63).astype(int)
        assert (g[:, 1:-1] == 63 + 8 * a).all(), a
        assert (g[:, 0] == 63 + 4 * a).all() and (g[:, -1] == 63 + 4 * a).all()


# ------------------------------ refinement costs: slices of full volumes
def geom_valid(base, valid, D, minD, maxD):
    """Label (y, x, d) of a refinement level exists iff its pixel has a
    parent, and the real disparity base + d is inside [minD, maxD] and
    projects into the right image (sgm_kernels.cu:1045-1054)."""
    h, w = base.shape
    real = base[:, :, None].astype(np.int64) + np.arange(D)[None, None, :]
    x = np.arange(w)[None, :, None]
    ok = (valid[:, :, None] != 0) & (real >= minD) & (real <= maxD) & (x - real >= 0) & (x - real < w)
    return ok, real


def local_case(seed, h=9, w=57, D=16, maxD=40, minD=0):
    r = rng(seed)
    left = r.integers(0, 256, (h, w), dtype=np.uint8)
    right = r.integers(0, 256, (h, w), dtype=np.uint8)
    base = np.clip(r.integers(-4, 5, (h, w)).cumsum(axis=1) + minD + 6, minD - 3,
                   maxD - 2).astype(np.int16)
    valid = (r.random((h, w)) > 0.15).astype(np.uint8)
    return left, right, base, valid


@pytest.mark.parametrize("minD,maxD,D", [(0, 40, 16), (0, 30, 32), (-5, 25, 16), (3, 44, 16)])
def test_census_cost_local_is_full_volume_slice(minD, maxD, D):
    """A refinement census cost is the full-range census cost volume read at
    the label's real disparity (sgm_kernels.cu:1112-1133 against :301-327);
    labels that do not exist are SHRT_MAX."""
    left, right, base, valid = local_case(31 + D - minD, minD=minD, maxD=maxD, D=D)
    cl, cr = O.census(left), O.census(right)
    full = O.census_cost(cl, cr, maxD - minD + 1, minD)          # index k <-> real minD + k
    ok, real = geom_valid(base, valid, D, minD, maxD)
    k = np.clip(real - minD, 0, maxD - minD)
    want = np.where(ok, np.take_along_axis(full, k, axis=2), MAX)
    got = O.census_cost_local(cl, cr, base, valid, D, maxD, minD)
    np.testing.assert_array_equal(got, want)
    assert (got[ok] <= 24).all() and (got[~ok] == MAX).all()


@pytest.mark.parametrize("block,cap,minD", [(1, 63, 0), (3, 63, 0), (5, 63, 0), (7, 63, 0),
                                            (5, 200, 0), (3, 63, -4), (5, 31, 2)])
def test_bt_window_cost_local_is_box_of_full_volume(block, cap, minD):
    """A refinement BT window cost is the clamped block x block window sum of
    the full-range BT volume at the label's real disparity: the two-pass box
    filter of bt_cost, sliced at base + d (sgm_kernels.cu:1075-1110 against
    :137-239 and :354-416; no saturation below 49 * 255)."""
    maxD = 40
    left, right, base, valid = local_case(41 + block + cap, minD=minD, maxD=maxD)
    D = 16
    full = O.box(O.bt_cost(left, right, maxD - minD + 1, minD, cap), block)
    ok, real = geom_valid(base, valid, D, minD, maxD)
    k = np.clip(real - minD, 0, maxD - minD)
    want = np.where(ok, np.take_along_axis(full, k, axis=2), MAX)
    got = O.bt_window_cost_local(left, right, base, valid, D, maxD, block, minD, cap)
    np.testing.assert_array_equal(got, want)


def test_bt_window_cost_local_caps_at_shrt_max_minus_one():
    """A large window of high BT costs overflows the int16 range: the local
    window cost is the plain sum capped at SHRT_MAX - 1 (sgm_kernels.cu:1105-1108;
    SHRT_MAX itself marks a label that does not exist).  The sums are formed
    here in int64 from the per-pixel BT volume, window by window."""
    r = rng(77)
    h, w, D, maxD, block, cap = 9, 40, 16, 24, 15, 63
    left = (r.integers(0, 2, (h, w)) * 255).astype(np.uint8)
    right = (r.integers(0, 2, (h, w)) * 255).astype(np.uint8)
    base = np.clip(r.integers(-2, 3, (h, w)).cumsum(axis=1) + 4, 0, maxD - D + 1).astype(np.int16)
    valid = np.ones((h, w), np.uint8)
    full = O.bt_cost(left, right, maxD + 1, 0, cap).astype(np.int64)   # [y][x][real]
    half = block // 2
    ok, real = geom_valid(base, valid, D, 0, maxD)
    k = np.clip(real, 0, maxD)
    sums = np.zeros((h, w, D), np.int64)
    ys, xs = np.arange(h), np.arange(w)
    for dy in range(-half, half + 1):
        yy = np.clip(ys + dy, 0, h - 1)
        for dx in range(-half, half + 1):
            xx = np.clip(xs + dx, 0, w - 1)
            sums += np.take_along_axis(full[yy][:, xx], k, axis=2)
    want = np.where(ok, np.minimum(sums, MAX - 1), MAX)
    got = O.bt_window_cost_local(left, right, base, valid, D, maxD, block, 0, cap)
    assert (sums[ok] > MAX - 1).any()          # the case reaches the cap
    np.testing.assert_array_equal(got, want)


def test_pyramid_rl_wta_three_geometric_labels():
    """Near the left image edge a pixel has only x + 1 labels that project into
    the right image; with exactly three (labels 0..2) a best at 0 and an
    alternative at 2 make a valid pixel (validCount >= 3, sgm_kernels.cu:1420)."""
    h, w, D, maxD = 1, 30, 16, 20
    base = np.zeros((h, w), np.int16)
    valid = np.ones((h, w), np.uint8)
    C = np.full((h, w, D), 20, np.int16)
    C[..., 0] = 0
    C[..., 2] = 10
    img = np.zeros((h, w), np.uint8)
    out = O.path_rl_pyr_wta(C, base, valid, img, 0, 0, maxD, uniq=0, min_conf=0.0)
    assert out.disp[0, 2] == 0 and out.conf[0, 2] > 0    # x = 2: labels 0, 1, 2 project
    assert out.disp[0, 1] == -16                          # x = 1: two labels, no alternative


# -------------------------- pyramid recurrences: real-disparity textbook
def textbook_pyramid(C, base, valid, P1, P2, minD, maxD, direction, lut=None, img=None):
    """Pyramid path costs along one direction ('lr', 'rl' or 'tb') in
    real-disparity space.  A label exists iff its geometry is valid and its
    cost is below MAX; every transition from the previous pixel's labels is
    priced by |real - real'| (0, P1, P2'); a pixel whose previous pixel had no
    path starts over at L = C.  P2' = lut[|I(p) - I(prev)|] when given."""
    h, w, D = C.shape
    ok, real = geom_valid(base, valid, D, minD, maxD)
    ok &= C < MAX
    out = np.full((h, w, D), MAX, np.int64)
    if direction == "tb":
        lines = [[(y, x) for y in range(h)] for x in range(w)]
    elif direction == "lr":
        lines = [[(y, x) for x in range(w)] for y in range(h)]
    else:
        lines = [[(y, x) for x in range(w - 1, -1, -1)] for y in range(h)]
    for line in lines:
        prev = None          # {real: L} of the previous pixel's labels with a path
        prev_p = None
        for (y, x) in line:
            L = np.full(D, MAX, np.int64)
            if not ok[y, x].any():
                out[y, x] = L
                prev, prev_p = None, (y, x)
                continue
            if prev is None:
                L[ok[y, x]] = C[y, x][ok[y, x]]
            else:
                p2 = P2 if lut is None else int(lut[abs(int(img[y, x]) - int(img[prev_p]))])
                mprev = min(prev.values())
                for d in range(D):
                    if not ok[y, x, d]:
                        continue
                    best = min(mprev + p2, MAX)
                    for rp, lv in prev.items():
                        dr = abs(rp - int(real[y, x, d]))
                        best = min(best, lv + (0 if dr == 0 else P1 if dr == 1 else p2))
                    L[d] = MAX if best >= MAX else min(MAX, int(C[y, x, d]) + best - mprev)
            out[y, x] = L
            cur = {int(real[y, x, d]): int(L[d]) for d in range(D) if L[d] < MAX}
            prev = cur if cur else None
            prev_p = (y, x)
    return out


def pyramid_case(seed, h=6, w=44, D=16, maxD=34):
    r = rng(seed)
    C = r.integers(0, 25, (h, w, D)).astype(np.int16)
    base = np.clip(r.integers(-3, 4, (h, w)).cumsum(axis=1) + 5, 0, 22).astype(np.int16)
    valid = (r.random((h, w)) > 0.1).astype(np.uint8)
    C[r.random((h, w, D)) < 0.05] = MAX
    img = r.integers(0, 256, (h, w), dtype=np.uint8)
    return C, base, valid, img


@pytest.mark.parametrize("adaptive", [False, True])
def test_pyramid_tb_matches_textbook_and_accumulates(adaptive):
    """sgm_kernels.cu:1263-1334: the T->B recurrence over columns, added to
    the L->R result with the pyramid saturation: MAX if either is MAX, else
    min(MAX - 1, old + L)."""
    C, base, valid, img = pyramid_case(51 + adaptive)
    maxD = 34
    lut = O.p2_lut(12, 1, 2.0, 10.0, 0.6, 1.8)[0] if adaptive else None
    acc0 = rng(52).integers(0, 30000, C.shape).astype(np.int16)
    acc0[rng(53).random(C.shape) < 0.05] = MAX
    got = O.path_pyr("tb", C, base, valid, img, 3, 12, maxD, acc=acc0, lut=lut)
    L = textbook_pyramid(C, base, valid, 3, 12, 0, maxD, "tb", lut, img)
    want = np.where((acc0 < MAX) & (L < MAX), np.minimum(MAX - 1, acc0.astype(np.int64) + L), MAX)
    np.testing.assert_array_equal(got, want)


def test_pyramid_lr_matches_textbook_adaptive():
    C, base, valid, img = pyramid_case(54)
    lut = O.p2_lut(12, 1, 2.0, 10.0, 0.6, 1.8)[0]
    got = O.path_pyr("lr", C, base, valid, img, 3, 12, 34, lut=lut)
    np.testing.assert_array_equal(got, textbook_pyramid(C, base, valid, 3, 12, 0, 34, "lr", lut, img))


def pyramid_wta_numpy(S, lv, base, valid, D, minD, maxD, uniq, min_conf):
    """Per-pixel selection of the pyramid R->L kernel (sgm_kernels.cu:1412-1478)
    restated: argmin with ties to the smallest label; second best over existing
    labels more than one label away; valid only with >= 3 geometric labels,
    finite best and alternative, uniqueness, no artificial window edge (best at
    label 0 while the window could start lower, or at label D-1 while it could
    reach higher), and confidence >= min_conf; 1/16-px parabola when both
    neighbours are finite; right-view projection keeps the lowest cost, and
    among equal costs the pixel visited first (largest x)."""
    h, w = S.shape[:2]
    disp = np.full((h, w), (minD - 1) * 16, np.int16)
    conf = np.zeros((h, w), np.float32)
    d2 = np.full((h, w), (minD - 1) * 16, np.int16)
    d2c = np.full((h, w), MAX, np.int64)
    geo, _ = geom_valid(base, valid, D, minD, maxD)
    for y in range(h):
        for x in range(w - 1, -1, -1):
            s = S[y, x]
            vc = int(geo[y, x].sum())
            b = int(s.argmin())
            m = int(s[b])
            alts = [int(s[d]) for d in range(D) if lv[y, x, d] and abs(d - b) > 1]
            alt = min(alts) if alts else MAX
            okp = vc >= 3 and m < MAX and alt < MAX
            if okp and uniq > 0 and alt * (100 - uniq) < m * 100:
                okp = False
            bs = int(base[y, x])
            if (b == 0 and bs > minD) or (b == D - 1 and bs + D - 1 < min(maxD, x)):
                okp = False
            cf = np.float32(0)
            if okp:
                cf = np.float32(np.float32(alt) - np.float32(m)) / np.float32(max(m, 1))
                cf = np.float32(min(max(cf, np.float32(0)), np.float32(1)))
                if cf < np.float32(min_conf):
                    okp = False
            if not okp:
                continue
            v = (bs + b) * 16
            if 0 < b < D - 1 and max(s[b - 1], s[b], s[b + 1]) < MAX:
                c0, c1, c2 = int(s[b - 1]), int(s[b]), int(s[b + 1])
                den = max(c0 + c2 - 2 * c1, 1)
                num = (c0 - c2) * 16 + den
                v += abs(num) // (2 * den) * (1 if num >= 0 else -1)
            disp[y, x], conf[y, x] = v, cf
            x2 = x - (bs + b)
            if 0 <= x2 < w and d2c[y, x2] > m:
                d2c[y, x2], d2[y, x2] = m, bs + b
    return disp, conf, d2, d2c.astype(np.int16)


@pytest.mark.parametrize("uniq,min_conf,accumulate", [(10, 0.05, True), (0, 0.0, True),
                                                      (25, 0.2, True), (10, 0.05, False)])
def test_pyramid_rl_wta_brute_force(uniq, min_conf, accumulate):
    """sgm_kernels.cu:1336-1478 on small volumes: the R->L recurrence from the
    real-disparity textbook, the pyramid sum rule, then every selection rule
    by brute force."""
    C, base, valid, img = pyramid_case(60 + uniq, h=7, w=48)
    D, maxD = 16, 34
    r = rng(61)
    acc0 = r.integers(0, 400, C.shape).astype(np.int16) if accumulate else None
    if accumulate:
        acc0[r.random(C.shape) < 0.03] = MAX
    out = O.path_rl_pyr_wta(C, base, valid, img, 3, 12, maxD, uniq=uniq, min_conf=min_conf,
                            acc=acc0)
    L = textbook_pyramid(C, base, valid, 3, 12, 0, maxD, "rl")
    geo, _ = geom_valid(base, valid, D, 0, maxD)
    lv = geo & (C < MAX)
    old = acc0.astype(np.int64) if accumulate else np.zeros(C.shape, np.int64)
    S = np.where(lv & (L < MAX) & (old < MAX), np.minimum(MAX - 1, old + L), MAX)
    np.testing.assert_array_equal(out.acc, S)
    disp, conf, d2, d2c = pyramid_wta_numpy(S, lv, base, valid, D, 0, maxD, uniq, min_conf)
    np.testing.assert_array_equal(out.disp, disp)
    np.testing.assert_array_equal(out.conf, conf)
    np.testing.assert_array_equal(out.disp2, d2)
    np.testing.assert_array_equal(out.disp2_cost, d2c)
    assert (disp != -16).sum() > 20          # the case exercises accepted pixels


def test_pyramid_rl_wta_artificial_edges():
    """A window whose cheapest label is its first (or last) label while the
    window could extend that way is rejected; at the true range ends it is
    kept (sgm_kernels.cu:1434-1440)."""
    h, w, D, maxD = 1, 60, 16, 40
    base = np.full((h, w), 10, np.int16)
    valid = np.ones((h, w), np.uint8)
    C = np.full((h, w, D), 20, np.int16)
    C[..., 0] = 0                                  # best at label 0 everywhere
    img = np.zeros((h, w), np.uint8)
    out = O.path_rl_pyr_wta(C, base, valid, img, 0, 0, maxD, uniq=0, min_conf=0.0)
    assert (out.disp[0, 25:] == -16).all()        # base 10 > minD 0: artificial left edge
    base0 = np.zeros((h, w), np.int16)
    out = O.path_rl_pyr_wta(C, base0, valid, img, 0, 0, maxD, uniq=0, min_conf=0.0)
    assert (out.disp[0, 25:] == 0).all()          # base = minD: a true edge
    C2 = np.full((h, w, D), 20, np.int16)
    C2[..., D - 1] = 0                             # best at the last label
    out = O.path_rl_pyr_wta(C2, base, valid, img, 0, 0, maxD, uniq=0, min_conf=0.0)
    assert (out.disp[0, 41:] == -16).all()        # 25 < min(maxD, x): could reach higher
    base_top = np.full((h, w), maxD - D + 1, np.int16)
    out = O.path_rl_pyr_wta(C2, base_top, valid, img, 0, 0, maxD, uniq=0, min_conf=0.0)
    assert (out.disp[0, 41:] == maxD * 16).all()  # window ends at maxD: a true edge


# ---------------------------- single-scale R->L with penalties: textbook
@pytest.mark.parametrize("P1,P2,adaptive,accumulate", [(3, 11, False, False), (7, 40, True, True),
                                                       (4, 4, False, True)])
def test_single_scale_rl_matches_mirrored_textbook(P1, P2, adaptive, accumulate):
    """sgm_kernels.cu:665-735: the R->L path is the textbook recurrence on the
    mirrored row (start L = C at x = W-1), added to the accumulator with
    saturation at SHRT_MAX; WTA then follows the summed costs."""
    r = rng(70 + P2)
    h, w, D = 6, 52, 16
    C = r.integers(0, 60, (h, w, D)).astype(np.int16)
    img = r.integers(0, 256, (h, w), dtype=np.uint8)
    lut = O.p2_lut(P2, 1, 2.0, 10.0, 0.6, 1.8)[0] if adaptive else None
    acc0 = r.integers(0, 500, (h, w, D)).astype(np.int16) if accumulate else None
    out = O.path_rl_wta(C, img, P1, P2, uniq=10, acc=acc0, lut=lut)
    Lm = textbook_path(np.ascontiguousarray(C[:, ::-1]), P1, P2, lut,
                       np.ascontiguousarray(img[:, ::-1]))[:, ::-1]
    S = np.minimum(MAX, (acc0.astype(np.int64) if accumulate else 0) + Lm)
    np.testing.assert_array_equal(out.acc, S)
    # The selection on top of S, for pixels inside [D, W-1).
    for y in range(h):
        for x in range(D, w - 1):
            s = S[y, x]
            b, m = int(s.argmin()), int(s.min())
            alt = min(int(s[d]) for d in range(D) if abs(d - b) > 1)
            if alt * 90 < m * 100:
                assert out.disp[y, x] == -16
                continue
            want = b * 16
            if 0 < b < D - 1:
                c0, c1, c2 = (int(v) for v in s[b - 1:b + 2])
                den = max(c0 + c2 - 2 * c1, 1)
                num = (c0 - c2) * 16 + den
                want += abs(num) // (2 * den) * (1 if num >= 0 else -1)
            assert out.disp[y, x] == want, (y, x)
