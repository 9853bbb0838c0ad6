"""Oracle-anchored GPU tests at the benchmark workloads, the reference's own
acceptance scenes, and the parameter corners the smaller parity cases leave out.

* 640x480 D=64 BT single scale (P1 200, P2 800): C-ABI and StereoMatcher;
* 1920x1080 D=256 Census, three levels, adaptive P2, all three outputs;
* 32 x 960x540 D=128 Census three-level batch: every pair against the same pair
  alone, three pairs against the oracle;
* the 64-pair uneven batch (per-pair seeds, disparities up to each pair's D):
  every pair against the same pair alone, and the smallest pair of every
  (cost, depth, D) combination against the oracle;
* the serving stream's sizes up to 3840x2160 D=512 with four levels: one
  request of each size through the serving loop, against the oracle;
* tests/matcher_test.cpp:198-247 (640x480, pyramid_levels 0..4 x {BT,
  Census}, constant and slanted scenes): bit-exact with the oracle and the
  reference's thresholds on the GPU result; tests/pyramid_smoke.cpp:93-113;
* single scale at D=512 and D=1024, output subsets at pyramid levels, and
  Census byte volumes at minD < 0 and > 0.

Oracle runs go through a thread pool: ctypes releases the GIL during the call.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

from oracle import oracle as O
from pyrsgm_amd import synth
from tests import gpu_util as U
from tests.test_oracle_pins import disp_stats, pattern, sample_texture

pytestmark = pytest.mark.gpu

import pyrsgm_amd as P

POOL = ThreadPoolExecutor(max_workers=16)


@pytest.fixture(scope="module")
def ws():
    w = P.create_workspace()
    yield w
    P.destroy_workspace(w)


def params_of(spec):
    census = spec["cost"] == "census"
    return P.validate_params(P.Params(
        num_disparities=spec["D"], pyramid_levels=spec["levels"],
        cost_mode=P.CostMode.Census if census else P.CostMode.BirchfieldTomasi,
        p1=8 if census else 300, p2=32 if census else 2400,
        median_filter=spec.get("median", True), adaptive_p2=census))


def same_outputs(tag, got, want):
    U.assert_same(f"{tag} disp16", got[0], want[0])
    U.assert_same(f"{tag} disp32", got[1], want[1])
    U.assert_same(f"{tag} conf", got[2], want[2])


def oracle_outputs(o):
    assert o.rc == 0
    return o.disp16, o.disp32, o.conf


# ----------------------------------------------------- 640x480 D=64 BT
def test_640x480_d64_bt_single_scale(ws):
    pair = synth.config_pair("640x480_d64_bt_single")
    kw = dict(minD=0, block=5, P1=200, P2=800, uniq=10, disp12=1, cap=63, cost=0, post=3,
              paths=3, adaptive=0)
    print("case: 640x480 D=64 BT single scale, C-ABI", flush=True)
    rc, d16, d32, cf = U.single_scale(ws, pair.left, pair.right, 64, **kw)
    want = O.compute_single_scale(pair.left, pair.right, 64, **kw)
    assert rc == 0
    same_outputs("c-abi", (d16, d32, cf), oracle_outputs(want))
    valid = (d16 != -16).mean()
    assert valid > 0.5, valid
    print("case: 640x480 D=64 BT single scale, StereoMatcher", flush=True)
    p = P.validate_params(P.Params(num_disparities=64, pyramid_levels=0,
                                   cost_mode=P.CostMode.BirchfieldTomasi, p1=200, p2=800))
    got = U.matcher(pair.left, pair.right, p)
    same_outputs("matcher", got, oracle_outputs(O.matcher_compute(pair.left, pair.right, p)))


# ------------------------------------ 1920x1080 D=256 Census, 3 levels
def bench_1080p_params():
    return P.validate_params(P.Params(
        num_disparities=256, pyramid_levels=2, cost_mode=P.CostMode.Census, p1=8, p2=32,
        uniqueness_ratio=10, disp12_max_diff=1, lr_check=True, median_filter=True,
        adaptive_p2=True, pyramid_coarse_disparities=128))


def test_1920x1080_d256_census_three_levels():
    pair = synth.config_pair("1920x1080_d256_census_l2")
    p = bench_1080p_params()
    fut = POOL.submit(O.matcher_compute, pair.left, pair.right, p)
    print("case: 1920x1080 D=256 census 3 levels", flush=True)
    got = U.matcher(pair.left, pair.right, p)
    same_outputs("1080p", got, oracle_outputs(fut.result()))
    # The scene's planes are recovered: most valid pixels within 1 px.
    valid = got[1] > 0
    err = np.abs(got[1] - pair.gt)[valid]
    assert valid.mean() > 0.5 and (err <= 1).mean() > 0.8, (valid.mean(), (err <= 1).mean())


# ------------------------------------ 32 x 960x540 D=128 Census batch
def test_batch_32x960x540_d128():
    n, w, h = 32, 960, 540
    pairs = [synth.config_pair("960x540_d128_census_l2", i) for i in range(n)]
    p = P.validate_params(P.Params(num_disparities=128, pyramid_levels=2,
                                   cost_mode=P.CostMode.Census, p1=8, p2=32, adaptive_p2=True))
    futs = {i: POOL.submit(O.matcher_compute, pairs[i].left, pairs[i].right, p) for i in (0, 13, 31)}
    L = torch.from_numpy(np.stack([q.left for q in pairs])).cuda()
    R = torch.from_numpy(np.stack([q.right for q in pairs])).cuda()
    d16 = torch.full((n, h, w), 1234, dtype=torch.int16, device="cuda")
    d32 = torch.full((n, h, w), -7.0, dtype=torch.float32, device="cuda")
    cf = torch.full((n, h, w), -7.0, dtype=torch.float32, device="cuda")
    bm = P.BatchMatcher(w, h, n, p)
    print("case: batch 32x960x540 D=128, graph replay", flush=True)
    bm.capture(L, R, d16, d32, cf)
    bm.launch()
    torch.cuda.synchronize()
    bm.close()
    got = [(d16[i].cpu().numpy(), d32[i].cpu().numpy(), cf[i].cpu().numpy()) for i in range(n)]
    m = P.StereoMatcher(w, h, p)
    a16 = torch.empty((h, w), dtype=torch.int16, device="cuda")
    a32 = torch.empty((h, w), dtype=torch.float32, device="cuda")
    ac = torch.empty((h, w), dtype=torch.float32, device="cuda")
    for i in range(n):
        m.compute(L[i], R[i], a16, a32, ac)
        torch.cuda.synchronize()
        same_outputs(f"pair {i} batch vs alone", got[i],
                     (a16.cpu().numpy(), a32.cpu().numpy(), ac.cpu().numpy()))
    m.close()
    for i, f in futs.items():
        same_outputs(f"pair {i} vs oracle", got[i], oracle_outputs(f.result()))


# ------------------------------------------------- 64-pair uneven batch
def test_uneven_64_pairs():
    specs = [synth.uneven_spec(i) for i in range(64)]
    pairs = synth.spec_pairs(specs)
    # The smallest pair of each (cost, depth, D) combination goes to the oracle.
    pick = {}
    for i, s in enumerate(specs):
        k = (s["cost"], s["levels"], s["D"])
        if k not in pick or s["w"] * s["h"] < specs[pick[k]]["w"] * specs[pick[k]]["h"]:
            pick[k] = i
    params = [params_of(s) for s in specs]
    futs = {i: POOL.submit(O.matcher_compute, pairs[i].left, pairs[i].right, params[i])
            for i in sorted(pick.values())}
    io = []
    for s, q in zip(specs, pairs):
        h, w = s["h"], s["w"]
        io.append((torch.from_numpy(q.left).cuda(), torch.from_numpy(q.right).cuda(),
                   torch.full((h, w), 1234, dtype=torch.int16, device="cuda"),
                   torch.full((h, w), -7.0, dtype=torch.float32, device="cuda"),
                   torch.full((h, w), -7.0, dtype=torch.float32, device="cuda")))
    mb = P.MixedBatch([(s["w"], s["h"], p) for s, p in zip(specs, params)])
    print("case: uneven 64 pairs, mixed batch graph", flush=True)
    mb.capture(io)
    mb.launch()
    torch.cuda.synchronize()
    mb.close()
    got = [tuple(t.cpu().numpy() for t in x[2:]) for x in io]
    del io
    for i, (s, q) in enumerate(zip(specs, pairs)):
        print(f"case: uneven pair {i} {s}", flush=True)
        alone = U.matcher(q.left, q.right, params[i])
        same_outputs(f"pair {i} mixed vs alone", got[i], alone)
    for i, f in futs.items():
        same_outputs(f"pair {i} {specs[i]} vs oracle", got[i], oracle_outputs(f.result()))
    assert len(futs) >= 20


# --------------------------------------- serving sizes up to 4K D=512
def test_serving_sizes_against_oracle():
    specs = []
    for k, (w, h, D, lv) in enumerate(synth.SERVING_SIZES):
        specs.append({"w": w, "h": h, "D": D, "levels": lv, "cost": ("bt", "census")[k % 2],
                      "median": True, "seed": 3900 + k})
    specs.append(dict(specs[-1], cost="bt", seed=3950))  # 4K with both costs
    pairs = synth.spec_pairs(specs)
    params = [params_of(s) for s in specs]
    futs = [POOL.submit(O.matcher_compute, q.left, q.right, p) for q, p in zip(pairs, params)]
    srv = P.Server(max_batch=4)
    outs, tickets = [], []
    print("case: serving one request per size, 640x480 .. 3840x2160 D=512", flush=True)
    for s, q, p in zip(specs, pairs, params):
        h, w = s["h"], s["w"]
        o = (torch.full((h, w), 1234, dtype=torch.int16, device="cuda"),
             torch.full((h, w), -7.0, dtype=torch.float32, device="cuda"),
             torch.full((h, w), -7.0, dtype=torch.float32, device="cuda"))
        L, R = torch.from_numpy(q.left).cuda(), torch.from_numpy(q.right).cuda()
        tickets.append(srv.submit(L, R, p, disp16=o[0], disp32=o[1], conf=o[2]))
        outs.append((L, R) + o)
    res = [srv.wait(t) for t in tickets]
    srv.close()
    assert all(r.status == 0 for r in res)
    for i, (s, f) in enumerate(zip(specs, futs)):
        got = tuple(t.cpu().numpy() for t in outs[i][2:])
        same_outputs(f"request {s}", got, oracle_outputs(f.result()))


# ------------------------ reference acceptance: matcher_test.cpp:198-247
def constant_scene(w=640, h=480, d=40):
    return pattern(w, h, d, False), pattern(w, h, d, True), np.full((h, w), float(d), np.float32)


def slanted_scene(w=640, h=480, d0=12.0, slope=0.06):
    ys = np.arange(h)[:, None] * np.ones((1, w), np.int64)
    xs = np.arange(w, dtype=np.float32)[None, :] * np.ones((h, 1), np.float32)
    left = sample_texture(xs, ys)
    right = sample_texture((xs + np.float32(d0)) / np.float32(1.0 - slope), ys)
    return left, right, np.float32(d0) + np.float32(slope) * xs


@pytest.mark.parametrize("levels", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("census", [False, True])
def test_reference_acceptance_scenes(levels, census):
    p = P.validate_params(P.Params(num_disparities=64, pyramid_levels=levels,
                                   cost_mode=P.CostMode(int(census)),
                                   p1=8 if census else 300, p2=32 if census else 2400))
    const, slant = constant_scene(), slanted_scene()
    fc = POOL.submit(O.matcher_compute, const[0], const[1], p)
    fs = POOL.submit(O.matcher_compute, slant[0], slant[1], p)
    print(f"case: acceptance levels={levels} census={census}", flush=True)
    gc = U.matcher(const[0], const[1], p)
    gs = U.matcher(slant[0], slant[1], p)
    same_outputs("constant", gc, oracle_outputs(fc.result()))
    same_outputs("slanted", gs, oracle_outputs(fs.result()))
    # The reference's thresholds, on the GPU's own outputs.
    dens, good, _ = disp_stats(gc[1], const[2], 64 + 8)
    assert dens > 0.5 and good > 0.95, (dens, good)
    inv = gc[0] == -16
    assert inv.any() and (gc[1][inv] == 0).all()
    ok = ~inv & (gc[1] != 0)
    assert np.abs(gc[0][ok] / 16.0 - gc[1][ok]).max() <= 1e-6
    _, good, mae = disp_stats(gs[1], slant[2], 64 + 8)
    assert good > 0.9 and mae < 0.5, (good, mae)


def test_reference_pyramid_smoke_chain(ws):
    """tests/pyramid_smoke.cpp:93-113 through the C-ABI: each level refines
    the GPU's own parent result, compared with the oracle on that parent."""
    sizes = [(40, 16, 2), (80, 32, 4), (160, 64, 8), (320, 128, 16), (640, 256, 32)]
    maxd = [31, 63, 127, 255, 511]
    kw = dict(block=3, P1=24, P2=96, uniq=0, disp12=1, cap=63, cost=0, post=1, paths=3,
              min_conf=0.0, adaptive=1, alpha=2.0, beta=14.0, lower=0.75, upper=2.0)
    parent = None
    for i, (w, h, d) in enumerate(sizes):
        l, r = pattern(w, h, d, False), pattern(w, h, d, True)
        print(f"case: pyramid smoke level {i} {w}x{h}", flush=True)
        if i == 0:
            rc, d16, d32, cf = U.level(ws, l, r, 32, maxd[i], refine=False, **kw)
            want = O.compute_level(l, r, 32, maxd[i], refine=False, **kw)
        else:
            rc, d16, d32, cf = U.level(ws, l, r, 16, maxd[i], parent[0], parent[1],
                                       refine=True, **kw)
            want = O.compute_level(l, r, 16, maxd[i], parent[0], parent[1], refine=True, **kw)
        assert rc == 0
        same_outputs(f"level {i}", (d16, d32, cf), oracle_outputs(want))
        v = d32[d32 > 0]
        tol = 2.0 if i == 4 else 1.5
        assert v.size > 0 and (np.abs(v - d) <= tol).sum() >= v.size / 2
        parent = (d16, cf)


# ------------------------------------------ single scale at D=512 / 1024
@pytest.mark.parametrize("w,h,D,minD,cost", [(700, 24, 512, 0, 0), (700, 24, 512, -3, 1),
                                             (1200, 12, 1024, 0, 1), (1200, 12, 1024, 5, 0)])
def test_single_scale_large_d(ws, w, h, D, minD, cost):
    left, right = (lambda q: (q.left, q.right))(synth.make_pair(w, h, D - 40, seed=D + w))
    P1, P2 = (8, 32) if cost else (300, 2400)
    kw = dict(minD=minD, block=5, P1=P1, P2=P2, uniq=10, disp12=1, cap=63, cost=cost, post=3,
              paths=3, adaptive=cost)
    fut = POOL.submit(O.compute_single_scale, left, right, D, **kw)
    print(f"case: single scale {w}x{h} D={D} minD={minD} cost={cost}", flush=True)
    rc, d16, d32, cf = U.single_scale(ws, left, right, D, **kw)
    assert rc == 0
    same_outputs(f"D={D}", (d16, d32, cf), oracle_outputs(fut.result()))


# ------------------------------------ output subsets at pyramid levels
SUBSETS = [(True, False, False), (False, True, False), (False, False, True),
           (True, False, True), (False, True, True), (True, True, False)]


@pytest.mark.parametrize("levels", [1, 2])
def test_matcher_output_subsets_pyramid(levels):
    """StereoMatcher routing (stereo_matcher.cpp:248-253; sgm_kernels.cu:1837-1858):
    disp32 without conf uses an internal confidence; conf alone is the raw copy."""
    pair = synth.make_pair(256, 144, 100, seed=71 + levels)
    p = P.validate_params(P.Params(num_disparities=128, pyramid_levels=levels,
                                   cost_mode=P.CostMode.Census, p1=8, p2=32, adaptive_p2=True))
    for w16, w32, wc in SUBSETS:
        print(f"case: subset levels={levels} 16={w16} 32={w32} conf={wc}", flush=True)
        got = U.matcher(pair.left, pair.right, p, w16, w32, wc)
        want = O.matcher_compute(pair.left, pair.right, p, w16, w32, wc)
        assert want.rc == 0
        for name, g, o in zip(("disp16", "disp32", "conf"), got, oracle_outputs(want)):
            assert (g is None) == (o is None), name
            U.assert_same(f"{name} subset {(w16, w32, wc)}", g, o)


@pytest.mark.parametrize("refine", [False, True])
def test_level_output_subsets(ws, refine):
    """pyrsgm_compute_level routing at a coarse and a refinement level."""
    left, right = (lambda q: (q.left, q.right))(synth.make_pair(200, 96, 60, seed=81))
    if refine:
        pl, pr = O.pyrdown(left), O.pyrdown(right)
        parent = O.compute_level(pl, pr, 32, 31, block=3, P1=8, P2=32, cost=1, post=1,
                                 want32=False)
        args = (16, 63, parent.disp16, parent.conf)
    else:
        args = (64, 63)
    for w16, w32, wc in SUBSETS:
        kw = dict(block=3, P1=8, P2=32, cost=1, post=3, paths=3, adaptive=1, refine=refine,
                  min_conf=0.05, want16=w16, want32=w32, wantc=wc)
        rc, d16, d32, cf = U.level(ws, left, right, *args, **kw)
        want = O.compute_level(left, right, *args, **kw)
        assert rc == 0 and want.rc == 0
        same_outputs(f"subset {(w16, w32, wc)}", (d16, d32, cf), oracle_outputs(want))


# ------------------------ Census byte volumes away from minD = 0
@pytest.mark.parametrize("minD", [-8, 5])
def test_census_byte_coarse_nonzero_mind(ws, minD):
    left, right = (lambda q: (q.left, q.right))(synth.make_pair(176, 80, 50, seed=95 - minD))
    D, maxD = 64, minD + 63
    kw = dict(minD=minD, block=3, P1=8, P2=32, cost=1, post=1, paths=3, adaptive=1)
    rc, d16, d32, cf = U.level(ws, left, right, D, maxD, **kw)
    want = O.compute_level(left, right, D, maxD, **kw)
    assert rc == 0
    same_outputs(f"coarse minD={minD}", (d16, d32, cf), oracle_outputs(want))


@pytest.mark.parametrize("minD,maxD", [(-4, 70), (3, 80), (0, 10), (2, 9)])
def test_census_byte_refine_nonzero_mind_and_narrow(ws, minD, maxD):
    """Refinement census on byte volumes with minD < 0, minD > 0, and label
    ranges narrower than the 16-label window (maxD - minD + 1 < 16), which
    takes the staged form's fallback."""
    left, right = (lambda q: (q.left, q.right))(synth.make_pair(192, 72, maxD, seed=97 + maxD))
    pl, pr = O.pyrdown(left), O.pyrdown(right)
    pmaxD = max(minD + 1, maxD // 2)
    parent = O.compute_level(pl, pr, 64, pmaxD, minD=minD, block=3, P1=8, P2=32, cost=1,
                             post=1, want32=False)
    kw = dict(minD=minD, block=3, P1=8, P2=32, cost=1, post=3, paths=3, adaptive=1, refine=True,
              min_conf=0.05)
    rc, d16, d32, cf = U.level(ws, left, right, 16, maxD, parent.disp16, parent.conf, **kw)
    want = O.compute_level(left, right, 16, maxD, parent.disp16, parent.conf, **kw)
    assert rc == 0 and want.rc == 0
    same_outputs(f"refine minD={minD} maxD={maxD}", (d16, d32, cf), oracle_outputs(want))


def test_pyramid_rejects_nonzero_min_disparity():
    """The reference allows a pyramid only with min_disparity 0
    (stereo_matcher.cpp:100-148), so a StereoMatcher at minD != 0 exists only
    for pyramid_levels = 0."""
    with pytest.raises(ValueError):
        P.validate_params(P.Params(num_disparities=64, pyramid_levels=1, min_disparity=-4))
