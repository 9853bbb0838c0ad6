"""PFM / calib.txt / metrics / report helpers (host only, no GPU).

The library's functions are checked against independent numpy restatements of
the formats (PFM: "Pf", width height, scale whose sign gives the byte order,
rows bottom-up) and of the Middlebury metrics, and against small cases worked
out by hand.
"""
import math

import numpy as np
import pytest

from pyrsgm_amd import middlebury as M


def np_write_pfm(path, a, big_endian=False):
    a = np.asarray(a, np.float32)
    with open(path, "wb") as f:
        f.write(f"Pf\n{a.shape[1]} {a.shape[0]}\n{1.0 if big_endian else -1.0}\n".encode())
        f.write(np.flipud(a).astype(">f4" if big_endian else "<f4").tobytes())


def test_pfm_round_trip_and_layout(tmp_path):
    rng = np.random.default_rng(1)
    a = rng.normal(size=(7, 13)).astype(np.float32) * 50
    a[2, 3] = np.inf
    a[0, 0] = -0.0
    p = tmp_path / "d.pfm"
    M.write_pfm(p, a)
    ref = tmp_path / "ref.pfm"
    np_write_pfm(ref, a)
    assert p.read_bytes() == ref.read_bytes()
    b = M.read_pfm(p)
    assert b.shape == a.shape and b.dtype == np.float32
    assert np.array_equal(b.view(np.uint32), a.view(np.uint32))


def test_pfm_big_endian_and_bottom_up(tmp_path):
    a = np.arange(12, dtype=np.float32).reshape(3, 4) + 0.25
    p = tmp_path / "be.pfm"
    np_write_pfm(p, a, big_endian=True)
    assert np.array_equal(M.read_pfm(p), a)
    # The first stored row is the bottom image row.
    raw = p.read_bytes()
    first = np.frombuffer(raw[len(b"Pf\n4 3\n1.0\n"):][:16], ">f4")
    assert np.array_equal(first, a[2])


def test_pfm_rejections(tmp_path):
    p = tmp_path / "c.pfm"
    p.write_bytes(b"PF\n2 2\n-1.0\n" + bytes(48))
    with pytest.raises(ValueError, match="grayscale"):
        M.read_pfm(p)
    p.write_bytes(b"Pf\n4 4\n-1.0\n" + bytes(20))
    with pytest.raises(ValueError, match="truncated"):
        M.read_pfm(p)
    p.write_bytes(b"Pf\n0 4\n-1.0\n")
    with pytest.raises(ValueError, match="header"):
        M.read_pfm(p)
    with pytest.raises(ValueError, match="cannot open"):
        M.read_pfm(tmp_path / "missing.pfm")


def test_read_ndisp(tmp_path):
    c = tmp_path / "calib.txt"
    c.write_text("cam0=[3997.684 0 1176.728; 0 3997.684 1011.728; 0 0 1]\n"
                 "doffs=131.111\nbaseline=193.001\nwidth=2964\nheight=1988\nndisp=280\n"
                 "isint=0\nvmin=31\nvmax=257\n")
    assert M.read_ndisp(c) == 280
    c.write_text("width=10\n ndisp=5\n")  # must start the line
    assert M.read_ndisp(c) == 0
    assert M.read_ndisp(tmp_path / "none.txt") == 0


def test_disp16_to_float():
    d = np.array([[-16, 0, 1, 16], [17, 255 * 16, 32767, -32768]], np.int16)
    want = np.where(d > 0, d.astype(np.float32) / 16.0, 0.0).astype(np.float32)
    assert np.array_equal(M.disp16_to_float(d), want)


def np_evaluate(disp, gt, mask, max_disp):
    sel = np.isfinite(gt) & (gt > 0) & (gt <= max_disp)
    if mask is not None:
        sel &= mask == 255
    ev = int(sel.sum())
    ok = sel & (disp > 0)
    va = int(ok.sum())
    e = np.abs(disp[ok].astype(np.float64) - gt[ok].astype(np.float64))
    out = dict(evaluated=ev, valid=va)
    if va:
        out.update(avg_err=e.mean(), rms_err=math.sqrt((e * e).mean()),
                   bad05=100 * (e > 0.5).sum() / va, bad1=100 * (e > 1).sum() / va,
                   bad2=100 * (e > 2).sum() / va, bad4=100 * (e > 4).sum() / va)
    if ev:
        out.update(density=100 * va / ev, invalid=100 - 100 * va / ev,
                   bad1_all=100 * ((e > 1).sum() + ev - va) / ev,
                   bad2_all=100 * ((e > 2).sum() + ev - va) / ev)
    return out


def test_evaluate_matches_numpy():
    rng = np.random.default_rng(7)
    h, w = 40, 57
    gt = rng.uniform(0, 70, (h, w)).astype(np.float32)
    gt[rng.random((h, w)) < 0.1] = np.inf
    disp = (gt + rng.normal(0, 2, (h, w))).astype(np.float32)
    disp[rng.random((h, w)) < 0.15] = 0.0
    disp[~np.isfinite(disp)] = 0.0
    mask = np.where(rng.random((h, w)) < 0.8, 255, 128).astype(np.uint8)
    for mk, md in ((None, 1e30), (mask, 64.0)):
        got = M.evaluate(disp, gt, mk, md)
        want = np_evaluate(disp, gt, mk, md)
        for k, v in want.items():
            assert got[k] == pytest.approx(v, rel=1e-12, abs=1e-12), k


def test_evaluate_hand_case():
    # Four known pixels: errors 0.5 (not > 0.5), 1.5, invalid, and one unknown GT.
    gt = np.array([[10.0, 20.0], [30.0, np.inf]], np.float32)
    disp = np.array([[10.5, 18.5], [0.0, 5.0]], np.float32)
    m = M.evaluate(disp, gt)
    assert (m["evaluated"], m["valid"]) == (3, 2)
    assert m["avg_err"] == pytest.approx(1.0)
    assert m["rms_err"] == pytest.approx(math.sqrt((0.25 + 2.25) / 2))
    assert (m["bad05"], m["bad1"], m["bad2"]) == (50.0, 50.0, 0.0)
    assert m["density"] == pytest.approx(200 / 3)
    assert m["bad1_all"] == pytest.approx(200 / 3)  # one bad + one invalid of three


def test_reports():
    r1 = M.Result("s1", "pyrsgm", 640, 480, 64, gpu_ms=[1.0, 3.0, 2.0], wall_ms=[2.0, 4.0, 3.0],
                  metrics=M.evaluate(np.ones((2, 2), np.float32), np.ones((2, 2), np.float32)))
    r2 = M.Result("s2", "pyrsgm", 640, 480, 64, gpu_ms=[4.0], wall_ms=[5.0])
    r3 = M.Result("s3", "pyrsgm", 640, 480, 64, error="out of memory")
    csv = M.csv_rows([r1, r2, r3]).splitlines()
    assert csv[0] == ("scene,algo,width,height,D,gpu_ms_median,gpu_ms_mean,gpu_ms_min,"
                      "wall_ms_median,density_pct,bad0.5_pct,bad1_pct,bad2_pct,bad4_pct,"
                      "avg_err_px,rms_err_px,bad1_all_pct,bad2_all_pct,error")
    assert csv[1] == "s1,pyrsgm,640,480,64,2.000,2.000,1.000,3.000,100.00,0.00,0.00,0.00,0.00,0.000,0.000,0.00,0.00,"
    assert csv[2] == "s2,pyrsgm,640,480,64,4.000,4.000,4.000,5.000,,,,,,,,,,"
    assert csv[3].endswith(",out of memory")
    assert all(len(l.split(",")) == 19 for l in csv)
    md = M.markdown_summary([r1, r2, r3], {"pyrsgm": "census, 3 paths"}).splitlines()
    assert md[0].startswith("| algorithm | paths | avg GPU ms | FPS | wall ms | density %")
    # avg GPU ms = mean of per-scene medians (2, 4) = 3; FPS = 1000 / 3; metrics / 2 scenes.
    assert md[2] == "| pyrsgm | 3 | 3.00 | 333.3 | 4.00 | 50.0 | 0.00 | 0.00 | 0.00 | 0.00 | 0.00 |"
