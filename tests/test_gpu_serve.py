"""The serving loop: every result equals matching the pair alone, in any mix.

Requests of several sizes, parameter sets and output choices are submitted
from one thread at staggered times, so that they meet lanes that are idle,
busy or still forming a batch.  Each result must equal StereoMatcher on the
same pair, alone, bit for bit (int16 and float outputs); StereoMatcher's own
parity with the oracle is tested in test_gpu_parity.py.
"""
import time

import numpy as np
import pytest

from pyrsgm_amd import synth
from tests import gpu_util as U

pytestmark = pytest.mark.gpu

import torch

import pyrsgm_amd as P

PAD = 3  # row pitch != width for every request buffer


def dev(a):
    h, w = a.shape
    t = torch.zeros((h, w + PAD), dtype=torch.uint8, device="cuda")
    t[:, :w] = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t[:, :w]


def out(h, w, dtype, fill):
    return torch.full((h, w + PAD), fill, dtype=dtype, device="cuda")[:, :w]


# (w, h, D, levels, cost, adaptive, median, outputs)  outputs: 16 / 32 / c
KINDS = [
    (160, 96, 64, 1, 1, 1, 1, "16,32,c"),
    (200, 120, 64, 2, 1, 1, 1, "16,32,c"),
    (128, 80, 64, 0, 0, 0, 1, "16,32,c"),
    (176, 100, 64, 2, 0, 1, 0, "16"),
    (160, 96, 64, 1, 1, 1, 1, "16,c"),
    (144, 90, 64, 3, 1, 0, 1, "32"),
]


def params_of(D, levels, cost, adaptive, median):
    return P.Params(num_disparities=D, pyramid_levels=levels, cost_mode=P.CostMode(cost),
                    adaptive_p2=bool(adaptive), median_filter=bool(median),
                    p1=8 if cost else 200, p2=32 if cost else 800)


def make_requests(n, seed0):
    reqs = []
    for i in range(n):
        w, h, D, lv, cost, ad, med, outs = KINDS[i % len(KINDS)]
        p = synth.make_pair(w, h, D - 1, seed=seed0 + i)
        reqs.append(dict(pair=p, params=params_of(D, lv, cost, ad, med), outs=outs.split(",")))
    return reqs


def alone(req):
    o = req["outs"]
    return U.matcher(req["pair"].left, req["pair"].right, req["params"], want16="16" in o,
                     want32="32" in o, wantc="c" in o)


def serve_all(server, reqs, gap_s=0.0):
    """Submit every request (with a gap between submissions); return outputs by index."""
    bufs = []
    tickets = []
    for r in reqs:
        h, w = r["pair"].left.shape
        b = dict(left=dev(r["pair"].left), right=dev(r["pair"].right),
                 disp16=out(h, w, torch.int16, 1234) if "16" in r["outs"] else None,
                 disp32=out(h, w, torch.float32, -7.0) if "32" in r["outs"] else None,
                 conf=out(h, w, torch.float32, -7.0) if "c" in r["outs"] else None)
        bufs.append(b)
    torch.cuda.synchronize()  # inputs written before the server's streams read them
    for r, b in zip(reqs, bufs):
        tickets.append(server.submit(b["left"], b["right"], r["params"], disp16=b["disp16"],
                                     disp32=b["disp32"], conf=b["conf"]))
        if gap_s:
            time.sleep(gap_s)
    results = [server.wait(t) for t in tickets]
    for res in results:
        assert res.status == 0
        assert res.done_ns >= res.launch_ns >= res.arrival_ns
    got = []
    for b in bufs:
        got.append(tuple(None if b[k] is None else b[k].cpu().numpy() for k in ("disp16", "disp32", "conf")))
    return got, results


def check_same(reqs, got):
    for i, (r, g) in enumerate(zip(reqs, got)):
        want = alone(r)
        for name, gv, wv in zip(("disp16", "disp32", "conf"), g, want):
            U.assert_same(f"request {i} {name}", gv, wv)


def test_mixed_stream_equals_alone():
    reqs = make_requests(18, 5000)
    s = P.Server(max_batch=4)
    got, results = serve_all(s, reqs, gap_s=0.0005)
    st = s.stats()
    s.close()
    assert st.completed == len(reqs) and st.failed == 0
    check_same(reqs, got)


def test_burst_batches_and_equals_alone():
    # All requests at once: same-key requests share launches.
    reqs = make_requests(24, 6000)
    s = P.Server(max_batch=8, lanes_per_key=1)
    got, results = serve_all(s, reqs)
    st = s.stats()
    s.close()
    assert max(r.batch_size for r in results) > 1
    assert st.launches < len(reqs)
    check_same(reqs, got)


def test_eager_lanes_equal_graph_lanes():
    reqs = make_requests(12, 7000)
    s1 = P.Server(max_batch=4, use_graphs=False)
    g1, _ = serve_all(s1, reqs)
    s1.close()
    s2 = P.Server(max_batch=4, use_graphs=True)
    g2, _ = serve_all(s2, reqs)
    s2.close()
    for i, (a, b) in enumerate(zip(g1, g2)):
        for name, x, y in zip(("disp16", "disp32", "conf"), a, b):
            U.assert_same(f"request {i} {name}", x, y)


def test_small_budget_evicts_and_equals_alone():
    # A budget that holds about one lane: lanes are evicted and re-created,
    # requests wait for memory, and results are unchanged.
    reqs = make_requests(12, 8000)
    need = max(P.BatchMatcher.required_bytes(r["pair"].left.shape[1], r["pair"].left.shape[0], 1,
                                             r["params"]) for r in reqs)
    s = P.Server(memory_budget=int(need * 1.6), max_batch=2, lanes_per_key=1)
    got, _ = serve_all(s, reqs, gap_s=0.0002)
    st = s.stats()
    s.close()
    assert st.lanes_evicted > 0
    assert st.peak_bytes <= st.budget_bytes
    check_same(reqs, got)


def test_rejections():
    s = P.Server(memory_budget=1 << 20)
    L = torch.zeros((64, 96), dtype=torch.uint8, device="cuda")
    d16 = torch.zeros((64, 96), dtype=torch.int16, device="cuda")
    p = P.Params(num_disparities=64)
    with pytest.raises(ValueError, match="memory budget"):
        s.submit(L, L, p, disp16=d16)
    with pytest.raises(ValueError, match="no outputs"):
        s.submit(L, L, p)
    with pytest.raises(ValueError):
        s.submit(L, L, P.Params(p1=10, p2=5), disp16=d16)
    with pytest.raises(ValueError, match="unknown ticket"):
        s.wait(12345)
    s.close()


def test_wait_ends_when_another_call_takes_the_ticket():
    """wait() on a ticket that poll() retrieves first raises instead of
    blocking forever; exactly one of the two callers gets the result."""
    import threading
    p = synth.make_pair(160, 96, 63, seed=404)
    params = params_of(64, 1, 1, 1, 1)
    srv = P.Server(max_batch=4)
    try:
        for rep in range(6):
            L, R = dev(p.left), dev(p.right)
            o = out(96, 160, torch.int16, 1234)
            t = srv.submit(L, R, params, disp16=o)
            box = {}

            def waiter():
                try:
                    box["wait"] = srv.wait(t)
                except ValueError as e:
                    box["err"] = str(e)

            th = threading.Thread(target=waiter)
            th.start()
            polled = None
            deadline = time.time() + 60
            while th.is_alive() and polled is None and time.time() < deadline:
                polled = srv.poll(t)
            th.join(timeout=60)
            assert not th.is_alive(), "wait() still blocked after the ticket was retrieved"
            got = ("wait" in box) + (polled is not None)
            assert got == 1, (rep, box, polled)
            if polled is not None:
                assert "retrieved by another call" in box["err"]
            torch.cuda.synchronize()
    finally:
        srv.close()


def test_python_api_rejects_bad_tensors():
    """Shape, dtype, device and row-stride checks before any launch."""
    w, h = 96, 64
    params = P.validate_params(P.Params(num_disparities=32, pyramid_levels=1,
                                        cost_mode=P.CostMode.Census, p1=8, p2=32))
    m = P.StereoMatcher(w, h, params)
    L = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
    ok16 = torch.zeros((h, w), dtype=torch.int16, device="cuda")
    bad = [
        dict(left=torch.zeros((h - 1, w), dtype=torch.uint8, device="cuda")),   # too small
        dict(disp16=torch.zeros((h, w), dtype=torch.float32, device="cuda")),   # wrong dtype
        dict(disp16=torch.zeros((h, w), dtype=torch.int16)),                    # host tensor
        dict(disp16=torch.zeros((w, h), dtype=torch.int16, device="cuda").t()), # column-major
        dict(conf=torch.zeros((h, 2 * w), dtype=torch.float32, device="cuda")[:, ::2]),
    ]
    for kw in bad:
        args = dict(left=L, right=L, disp16=ok16, disp32=None, conf=None)
        args.update(kw)
        with pytest.raises(ValueError):
            m.compute(args["left"], args["right"], args["disp16"], args["disp32"], args["conf"])
    m.compute(L, L, ok16)          # a correct call still runs
    torch.cuda.synchronize()
    m.close()
    srv = P.Server(max_batch=2)
    with pytest.raises(ValueError):
        srv.submit(L, L, params, disp16=torch.zeros((h, w), dtype=torch.float32, device="cuda"))
    srv.close()
