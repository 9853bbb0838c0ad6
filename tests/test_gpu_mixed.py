"""The mixed batch: pairs of different sizes, disparity ranges, pyramid depths,
cost modes, penalties and outputs in one fused pass must each equal the same
pair matched alone by StereoMatcher, bit for bit -- eagerly, under graph
replay, and under every tuning of the launch shape.

StereoMatcher's own parity with the oracle is tested in test_gpu_parity.py;
one pair here is also compared with the oracle directly.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pyrsgm_amd import synth
from tests import gpu_util as U

pytestmark = pytest.mark.gpu

import torch

import pyrsgm_amd as P

# (w, h, D, levels, cost, adaptive, median, P1, P2, outputs, paths, minD)
SPECS = [
    (200, 120, 64, 1, 1, 1, 1, 8, 32, "all", 3, 0),       # census, byte volumes
    (241, 135, 128, 2, 1, 1, 1, 8, 32, "all", 3, 0),      # census, 2 refinement levels
    (160, 90, 64, 0, 0, 0, 1, 300, 2400, "all", 3, 0),    # single-scale BT
    (192, 108, 64, 2, 0, 1, 0, 300, 2400, "d16", 3, 0),   # BT pyramid, adaptive
    (97, 61, 64, 3, 1, 0, 1, 8, 32, "d32", 3, 0),         # deep pyramid of a tiny image
    (176, 99, 64, 0, 1, 1, 0, 8, 32, "all", 2, 0),        # single-scale census, adaptive (image patch), 2 paths
    (150, 80, 48, 0, 0, 1, 1, 200, 800, "conf", 1, -8),   # single-scale BT, minD < 0, 1 path, adaptive
    (180, 100, 128, 1, 1, 1, 1, 8, 300, "all", 3, 0),     # census int16 volumes (P2 300)
    (200, 110, 64, 1, 1, 0, 0, 8, 16000, "all", 3, 0),    # exact path form (P2 16000)
    (320, 180, 128, 2, 1, 1, 1, 8, 32, "d16+conf", 3, 0),
    (200, 120, 64, 1, 1, 1, 1, 8, 32, "all", 3, 0),       # same key as pair 0, other scene
]


def params_of(s):
    w, h, D, levels, cost, adaptive, median, p1, p2, outs, paths, minD = s
    return P.validate_params(P.Params(
        num_disparities=D, min_disparity=minD, pyramid_levels=levels, cost_mode=P.CostMode(cost),
        p1=p1, p2=p2, adaptive_p2=bool(adaptive), median_filter=bool(median), num_paths=paths))


def wants(outs):
    if outs == "all":
        return True, True, True
    return "d16" in outs, "d32" in outs, "conf" in outs


@pytest.fixture(scope="module")
def scene():
    pairs, specs, refs = [], [], []
    for i, s in enumerate(SPECS):
        w, h, D = s[0], s[1], s[2]
        pr = synth.make_pair(w, h, D - 1, seed=7000 + i)
        prm = params_of(s)
        pairs.append(pr)
        specs.append((w, h, prm))
        refs.append(U.matcher(pr.left, pr.right, prm, *wants(s[9])))
    return pairs, specs, refs


def make_io(pairs):
    io = []
    for s, pr in zip(SPECS, pairs):
        h, w = pr.left.shape
        w16, w32, wc = wants(s[9])
        io.append((U.Img.upload(pr.left), U.Img.upload(pr.right),
                   U.Img(h, w, torch.int16, 1234) if w16 else None,
                   U.Img(h, w, torch.float32, -7.0) if w32 else None,
                   U.Img(h, w, torch.float32, -7.0) if wc else None))
    return io


def tensors(io):
    return [tuple(None if im is None else im.t[:, :im.w] for im in bufs) for bufs in io]


def reset(io):
    for bufs in io:
        for k, im in enumerate(bufs[2:]):
            if im is not None:
                im.t.fill_(1234 if k == 0 else -7.0)


def check(io, refs, what):
    for i, (bufs, ref) in enumerate(zip(io, refs)):
        for name, im, want in zip(("disp16", "disp32", "conf"), bufs[2:], ref):
            if want is not None:
                U.assert_same(f"{what}: pair {i} {name} {SPECS[i][:4]}", im.numpy(), want)


def test_mixed_eager_equals_alone(scene):
    pairs, specs, refs = scene
    mb = P.MixedBatch(specs)
    plan = mb.plan()
    pair_levels = sum(s[3] + 1 for s in SPECS)
    assert plan["path_tasks"] == pair_levels
    assert plan["queued_path_launches"] < pair_levels  # tasks of one kernel and step share a launch
    # Every other stage is queued too: no per-pair launch is left, and each pair
    # level contributes its descriptor pass, cost pass(es) and post as tasks,
    # plus one pyrDown task per pyramid level.
    assert plan["per_pair_stage_calls"] == 0
    pyr_levels = sum(s[3] for s in SPECS)
    assert plan["stage_tasks"] >= 3 * pair_levels + pyr_levels
    assert plan["queued_stage_launches"] < plan["stage_tasks"]
    io = make_io(pairs)
    print("case: mixed eager,", len(SPECS), "pairs, plan", plan, flush=True)
    mb.compute(tensors(io))
    U.sync()
    check(io, refs, "eager")
    ref = O.matcher_compute(pairs[1].left, pairs[1].right, specs[1][2])
    U.assert_same("pair 1 vs oracle", io[1][2].numpy(), ref.disp16)
    mb.close()


@pytest.mark.parametrize("branches,bpc,paired,queued", [(1, 0, -1, 1), (2, 1, 0, 0), (4, 2, 1, 1),
                                                       (3, 8, -1, 0), (4, 0, -1, 0)])
def test_mixed_tunings_equal_alone(scene, branches, bpc, paired, queued):
    pairs, specs, refs = scene
    mb = P.MixedBatch(specs)
    mb.set_tuning(branches, bpc, paired, queued)
    assert mb.tuning() == {"branches": branches, "blocks_per_cu": bpc, "paired": paired,
                           "queued_stages": queued}
    plan = mb.plan()
    assert (plan["per_pair_stage_calls"] == 0) == bool(queued)
    assert (plan["queued_stage_launches"] > 0) == bool(queued)
    io = make_io(pairs)
    print("case: mixed tuning", branches, bpc, paired, queued, flush=True)
    for rep in range(2):
        reset(io)
        mb.compute(tensors(io))
        U.sync()
        check(io, refs, f"tuning {branches}/{bpc}/{paired}/{queued} run {rep}")
    mb.close()


@pytest.mark.parametrize("queued", [1, 0])
def test_mixed_graph_replays_equal_alone(scene, queued):
    pairs, specs, refs = scene
    mb = P.MixedBatch(specs)
    mb.set_tuning(queued_stages=queued)
    io = make_io(pairs)
    print("case: mixed graph, queued stages", queued, flush=True)
    mb.capture(tensors(io))
    junk = torch.zeros(1 << 20, device="cuda")
    for rep in range(3):
        reset(io)
        mb.launch()
        U.sync()
        check(io, refs, f"replay {rep}")
        # Eager work between replays, including an eager pass of the same batch.
        junk.add_(1.0)
        mb.compute(tensors(io))
        U.sync()
    mb.close()


def test_mixed_tune_keeps_results(scene):
    pairs, specs, refs = scene
    mb = P.MixedBatch(specs)
    io = make_io(pairs)
    print("case: mixed tune", flush=True)
    best, ms = mb.tune(tensors(io), reps=2)
    assert ms > 0
    assert mb.tuning() == best
    assert 1 <= best["branches"] <= 4 and best["paired"] in (-1, 0, 1)
    assert best["queued_stages"] in (0, 1)
    reset(io)
    mb.compute(tensors(io))
    U.sync()
    check(io, refs, f"tuned {best}")
    mb.close()


def test_mixed_rejections(scene):
    pairs, specs, _ = scene
    with pytest.raises(ValueError):
        P.MixedBatch([])
    with pytest.raises(ValueError):
        P.MixedBatch([(0, 80, P.Params(num_disparities=64))])  # empty image
    mb = P.MixedBatch(specs[:2])
    io = make_io(pairs[:2])
    t = tensors(io)
    with pytest.raises(ValueError):
        mb.compute(t[:1])
    with pytest.raises(RuntimeError):
        mb.compute([t[0], (t[1][0], t[1][1], None, None, None)])  # a pair with no output
    with pytest.raises(ValueError):
        mb.set_tuning(branches=5)
    with pytest.raises(ValueError):
        mb.set_tuning(queued_stages=2)
    with pytest.raises(RuntimeError):
        mb.launch()  # nothing captured
    mb.close()
