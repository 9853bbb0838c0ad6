"""ctypes front-end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/_build/liboracle.so`` (built by ``oracle/Makefile`` with gcc) and
exposes numpy-level calls.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module; the product
package never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PYRSGM_ORACLE_LIB points the tests at another build of the same source (a
# copy with a planted mistake, to check that the pins catch it).
_LIB_PATH = os.environ.get("PYRSGM_ORACLE_LIB") or os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

COST_BT, COST_CENSUS = 0, 1
POST_LRCHECK, POST_MEDIAN = 1, 2


def build() -> str:
    """Compile the oracle with gcc (used by __graft_entry__.build())."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(np.int16, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_ci, _cf, _sz, _vp = C.c_int, C.c_float, C.c_size_t, C.c_void_p


class OrcParams(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "num_disparities", "min_disparity", "block_size", "cost_mode", "p1", "p2",
        "uniqueness_ratio", "disp12_max_diff", "lr_check", "median_filter", "num_paths",
        "pre_filter_cap", "pyramid_levels", "pyramid_coarse_disparities",
        "pyramid_max_disparity")] + [
        ("pyramid_min_confidence", C.c_float), ("adaptive_p2", C.c_int),
        ("p2_alpha", C.c_float), ("p2_beta", C.c_float), ("p2_lower_mult", C.c_float),
        ("p2_upper_mult", C.c_float)]


class OrcLevelInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "level", "width", "height", "num_labels", "max_disparity", "block_size", "refine")]


def _declare(L):
    L.orc_pyrdown_u8.argtypes = [_u8p, _sz, _ci, _ci, _u8p, _sz, _ci, _ci]
    L.orc_census5x5.argtypes = [_u8p, _sz, _ci, _ci, _u32p]
    L.orc_census5x5.restype = None
    L.orc_bt_grad.argtypes = [_u8p, _sz, _ci, _ci, _ci, _u8p]
    L.orc_bt_grad.restype = None
    L.orc_bt_cost.argtypes = [_u8p, _u8p, _u8p, _sz, _u8p, _sz, _ci, _ci, _ci, _ci, _i16p]
    L.orc_bt_cost.restype = None
    for f in (L.orc_box_h, L.orc_box_v):
        f.argtypes = [_i16p, _i16p, _ci, _ci, _ci, _ci]
        f.restype = None
    L.orc_census_cost.argtypes = [_u32p, _u32p, _ci, _ci, _ci, _ci, _i16p]
    L.orc_census_cost.restype = None
    L.orc_p2_lut.argtypes = [_ci, _ci, _cf, _cf, _cf, _cf, _i32p]
    L.orc_path_lr.argtypes = [_i16p, _i16p, _u8p, _sz, _ci, _ci, _ci, _ci, _ci, _ci, _ci, _ci, _i32p]
    L.orc_path_lr.restype = None
    L.orc_path_tb.argtypes = [_i16p, _i16p, _u8p, _sz, _ci, _ci, _ci, _ci, _ci, _ci, _i32p]
    L.orc_path_tb.restype = None
    L.orc_path_rl_wta.argtypes = [_i16p, _i16p, _u8p, _sz, _i16p, _i16p, _i16p, _f32p,
                                  _ci, _ci, _ci, _ci, _ci, _ci, _ci, _ci, _ci, _i32p]
    L.orc_path_rl_wta.restype = None
    for f in (L.orc_pseudo_lr_check, L.orc_strict_lr_check):
        f.argtypes = [_i16p, _i16p, _f32p, _ci, _ci, _ci, _ci]
        f.restype = None
    L.orc_median3x3.argtypes = [_i16p, _i16p, _ci, _ci]
    L.orc_median3x3.restype = None
    L.orc_prepare_refine_labels.argtypes = [_i16p, _f32p, _ci, _ci, _sz, _sz, _i16p, _u8p,
                                            _ci, _ci, _ci, _ci, _ci, _cf]
    L.orc_prepare_refine_labels.restype = None
    L.orc_census_cost_local.argtypes = [_u32p, _u32p, _i16p, _u8p, _i16p, _ci, _ci, _ci, _ci, _ci]
    L.orc_census_cost_local.restype = None
    L.orc_bt_window_cost_local.argtypes = [_u8p, _u8p, _u8p, _sz, _u8p, _sz, _i16p, _u8p, _i16p,
                                           _ci, _ci, _ci, _ci, _ci, _ci]
    L.orc_bt_window_cost_local.restype = None
    for f in (L.orc_path_lr_pyr, L.orc_path_tb_pyr):
        f.argtypes = [_i16p, _i16p, _i16p, _u8p, _u8p, _sz, _ci, _ci, _ci, _ci, _ci,
                      _ci, _ci, _ci, _i32p]
        f.restype = None
    L.orc_path_rl_pyr_wta.argtypes = [_i16p, _i16p, _i16p, _u8p, _u8p, _sz, _i16p, _i16p, _i16p,
                                      _f32p, _ci, _ci, _ci, _ci, _ci, _ci, _ci, _ci, _cf, _ci,
                                      _ci, _i32p]
    L.orc_path_rl_pyr_wta.restype = None
    L.orc_compute_single_scale.argtypes = [_vp] * 5 + [_sz] * 5 + [_ci] * 14 + [_cf] * 4
    L.orc_compute_level.argtypes = [_vp] * 7 + [_sz] * 7 + [_ci] * 17 + [_cf, _ci] + [_cf] * 4
    L.orc_level_geometry.argtypes = [_ci, _ci, C.POINTER(OrcParams), C.POINTER(OrcLevelInfo)]
    L.orc_matcher_compute.argtypes = [_ci, _ci, C.POINTER(OrcParams), _vp, _vp, _vp, _vp, _vp]


def _ptr(a):
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------- stages
def pyrdown(src: np.ndarray) -> np.ndarray:
    h, w = src.shape
    dst = np.zeros(((h + 1) // 2, (w + 1) // 2), np.uint8)
    rc = lib().orc_pyrdown_u8(np.ascontiguousarray(src), w, w, h, dst, dst.shape[1],
                              dst.shape[1], dst.shape[0])
    if rc != 0:
        raise ValueError("pyrdown: bad size")
    return dst


def census(img: np.ndarray) -> np.ndarray:
    h, w = img.shape
    out = np.zeros((h, w), np.uint32)
    lib().orc_census5x5(np.ascontiguousarray(img), w, w, h, out)
    return out


def bt_grad(img: np.ndarray, cap: int = 63) -> np.ndarray:
    h, w = img.shape
    out = np.zeros((h, w), np.uint8)
    lib().orc_bt_grad(np.ascontiguousarray(img), w, w, h, cap, out)
    return out


def bt_cost(left, right, D, minD=0, cap=63):
    h, w = left.shape
    g1, g2 = bt_grad(left, cap), bt_grad(right, cap)
    out = np.zeros((h, w, D), np.int16)
    lib().orc_bt_cost(g1, g2, np.ascontiguousarray(left), w, np.ascontiguousarray(right), w,
                      w, h, D, minD, out)
    return out


def box(vol: np.ndarray, block: int) -> np.ndarray:
    h, w, D = vol.shape
    tmp = np.zeros_like(vol)
    out = np.zeros_like(vol)
    lib().orc_box_h(np.ascontiguousarray(vol), tmp, w, h, D, block)
    lib().orc_box_v(tmp, out, w, h, D, block)
    return out


def census_cost(cl, cr, D, minD=0):
    h, w = cl.shape
    out = np.zeros((h, w, D), np.int16)
    lib().orc_census_cost(np.ascontiguousarray(cl), np.ascontiguousarray(cr), w, h, D, minD, out)
    return out


def p2_lut(P2, adaptive=1, alpha=3.041, beta=16.137, lower=0.601, upper=1.77):
    lut = np.zeros(256, np.int32)
    on = lib().orc_p2_lut(P2, adaptive, alpha, beta, lower, upper, lut)
    return lut, bool(on)


def median3x3(d):
    h, w = d.shape
    out = np.zeros_like(d)
    lib().orc_median3x3(np.ascontiguousarray(d), out, w, h)
    return out


def path_lr(C_, img, P1, P2, startX=0, endX=None, lut=None):
    h, w, D = C_.shape
    acc = np.zeros_like(C_)
    lut_a = lut if lut is not None else np.zeros(256, np.int32)
    lib().orc_path_lr(acc, np.ascontiguousarray(C_), np.ascontiguousarray(img), w, w, h, D,
                      startX, w if endX is None else endX, P1, P2, int(lut is not None), lut_a)
    return acc


def path_tb(C_, img, P1, P2, acc=None, lut=None):
    h, w, D = C_.shape
    acc = np.zeros_like(C_) if acc is None else acc.copy()
    lut_a = lut if lut is not None else np.zeros(256, np.int32)
    lib().orc_path_tb(acc, np.ascontiguousarray(C_), np.ascontiguousarray(img), w, w, h, D,
                      P1, P2, int(lut is not None), lut_a)
    return acc


@dataclass
class WTAOut:
    acc: np.ndarray
    disp: np.ndarray
    disp2: np.ndarray
    disp2_cost: np.ndarray
    conf: np.ndarray


def path_rl_wta(C_, img, P1, P2, uniq=10, minD=0, acc=None, lut=None):
    h, w, D = C_.shape
    accumulate = acc is not None
    acc = np.zeros_like(C_) if acc is None else acc.copy()
    disp = np.full((h, w), (minD - 1) * 16, np.int16)
    d2 = np.full((h, w), (minD - 1) * 16, np.int16)
    d2c = np.full((h, w), 32767, np.int16)
    conf = np.zeros((h, w), np.float32)
    lut_a = lut if lut is not None else np.zeros(256, np.int32)
    lib().orc_path_rl_wta(acc, np.ascontiguousarray(C_), np.ascontiguousarray(img), w, disp, d2,
                          d2c, conf, w, h, D, minD, P1, P2, uniq, int(accumulate),
                          int(lut is not None), lut_a)
    return WTAOut(acc, disp, d2, d2c, conf)


def lr_check(disp, disp2, conf, minD=0, max_diff=1, strict=False):
    h, w = disp.shape
    d, c = disp.copy(), conf.copy()
    f = lib().orc_strict_lr_check if strict else lib().orc_pseudo_lr_check
    f(d, np.ascontiguousarray(disp2), c, w, h, minD, max_diff)
    return d, c


def prepare_refine_labels(pdisp, pconf, w, h, maxD, minD=0, localD=16, min_conf=0.05):
    ph, pw = pdisp.shape
    base = np.zeros((h, w), np.int16)
    valid = np.zeros((h, w), np.uint8)
    lib().orc_prepare_refine_labels(np.ascontiguousarray(pdisp), np.ascontiguousarray(pconf),
                                    pw, ph, 2 * pw, 4 * pw, base, valid, w, h, minD, maxD,
                                    localD, min_conf)
    return base, valid


def census_cost_local(cl, cr, base, valid, D, maxD, minD=0):
    h, w = cl.shape
    out = np.zeros((h, w, D), np.int16)
    lib().orc_census_cost_local(np.ascontiguousarray(cl), np.ascontiguousarray(cr),
                                np.ascontiguousarray(base), np.ascontiguousarray(valid), out,
                                w, h, D, minD, maxD)
    return out


def bt_window_cost_local(left, right, base, valid, D, maxD, block, minD=0, cap=63):
    h, w = left.shape
    g1, g2 = bt_grad(left, cap), bt_grad(right, cap)
    out = np.zeros((h, w, D), np.int16)
    lib().orc_bt_window_cost_local(g1, g2, np.ascontiguousarray(left), w,
                                   np.ascontiguousarray(right), w, np.ascontiguousarray(base),
                                   np.ascontiguousarray(valid), out, w, h, D, minD, maxD, block)
    return out


def path_pyr(kind, C_, base, valid, img, P1, P2, maxD, minD=0, acc=None, lut=None):
    """kind in {'lr','tb'}; returns the updated accumulator."""
    h, w, D = C_.shape
    acc = np.zeros_like(C_) if acc is None else acc.copy()
    lut_a = lut if lut is not None else np.zeros(256, np.int32)
    f = lib().orc_path_lr_pyr if kind == "lr" else lib().orc_path_tb_pyr
    f(acc, np.ascontiguousarray(C_), np.ascontiguousarray(base), np.ascontiguousarray(valid),
      np.ascontiguousarray(img), w, w, h, D, minD, maxD, P1, P2, int(lut is not None), lut_a)
    return acc


def path_rl_pyr_wta(C_, base, valid, img, P1, P2, maxD, uniq=10, min_conf=0.05, minD=0,
                    acc=None, lut=None):
    h, w, D = C_.shape
    accumulate = acc is not None
    acc = np.zeros_like(C_) if acc is None else acc.copy()
    disp = np.full((h, w), (minD - 1) * 16, np.int16)
    d2 = np.full((h, w), (minD - 1) * 16, np.int16)
    d2c = np.full((h, w), 32767, np.int16)
    conf = np.zeros((h, w), np.float32)
    lut_a = lut if lut is not None else np.zeros(256, np.int32)
    lib().orc_path_rl_pyr_wta(acc, np.ascontiguousarray(C_), np.ascontiguousarray(base),
                              np.ascontiguousarray(valid), np.ascontiguousarray(img), w, disp, d2,
                              d2c, conf, w, h, D, minD, maxD, P1, P2, uniq, min_conf,
                              int(accumulate), int(lut is not None), lut_a)
    return WTAOut(acc, disp, d2, d2c, conf)


# ------------------------------------------------------------ entry points
# Output buffers start from sentinels (1234 / -7.0) so that outputs an entry
# point leaves unwritten stay recognisable; the GPU tests fill theirs the same way.
FILL16, FILLF = 1234, -7.0


def _outputs(h, w, want16, want32, wantc):
    return (np.full((h, w), FILL16, np.int16) if want16 else None,
            np.full((h, w), FILLF, np.float32) if want32 else None,
            np.full((h, w), FILLF, np.float32) if wantc else None)


@dataclass
class Outputs:
    disp16: np.ndarray | None
    disp32: np.ndarray | None
    conf: np.ndarray | None
    rc: int


def compute_single_scale(left, right, D, minD=0, block=5, P1=300, P2=2400, uniq=10, disp12=1,
                         cap=63, cost=COST_BT, post=POST_LRCHECK | POST_MEDIAN, paths=3,
                         adaptive=0, alpha=3.041, beta=16.137, lower=0.601, upper=1.77,
                         want16=True, want32=True, wantc=True):
    h, w = left.shape
    d16, d32, cf = _outputs(h, w, want16, want32, wantc)
    l = np.ascontiguousarray(left, np.uint8)
    r = np.ascontiguousarray(right, np.uint8)
    rc = lib().orc_compute_single_scale(_ptr(l), _ptr(r), _ptr(d16), _ptr(d32), _ptr(cf),
                                        w, w, 2 * w, 4 * w, 4 * w, w, h, D, minD, block, P1, P2,
                                        uniq, disp12, cap, cost, post, paths, adaptive,
                                        alpha, beta, lower, upper)
    return Outputs(d16, d32, cf, rc)


def compute_level(left, right, D, maxD, parent_disp=None, parent_conf=None, minD=0, block=5,
                  P1=300, P2=2400, uniq=10, disp12=1, cap=63, cost=COST_BT, post=POST_LRCHECK,
                  paths=3, refine=False, min_conf=0.05, adaptive=0, alpha=3.041, beta=16.137,
                  lower=0.601, upper=1.77, want16=True, want32=True, wantc=True):
    h, w = left.shape
    d16, d32, cf = _outputs(h, w, want16, want32, wantc)
    l = np.ascontiguousarray(left, np.uint8)
    r = np.ascontiguousarray(right, np.uint8)
    pd = None if parent_disp is None else np.ascontiguousarray(parent_disp, np.int16)
    pc = None if parent_conf is None else np.ascontiguousarray(parent_conf, np.float32)
    ph, pw = (0, 0) if pd is None else pd.shape
    rc = lib().orc_compute_level(_ptr(l), _ptr(r), _ptr(pd), _ptr(pc), _ptr(d16), _ptr(d32),
                                 _ptr(cf), w, w, 2 * pw, 4 * pw, 2 * w, 4 * w, 4 * w, w, h, pw,
                                 ph, D, minD, maxD, block, P1, P2, uniq, disp12, cap, cost, post,
                                 paths, int(refine), min_conf, adaptive, alpha, beta, lower,
                                 upper)
    return Outputs(d16, d32, cf, rc)


def params_struct(p) -> OrcParams:
    """p: pyrsgm_amd.Params (already validated) or any object with the same fields."""
    s = OrcParams()
    for name, _ in OrcParams._fields_:
        v = getattr(p, name)
        if name == "cost_mode":
            v = int(v)
        setattr(s, name, type(getattr(s, name))(v) if not isinstance(v, bool) else int(v))
    return s


def level_geometry(w, h, p):
    s = params_struct(p)
    info = (OrcLevelInfo * 5)()
    n = lib().orc_level_geometry(w, h, C.byref(s), info)
    return [dict((f, getattr(info[i], f)) for f, _ in OrcLevelInfo._fields_) for i in range(n)]


def matcher_compute(left, right, p, want16=True, want32=True, wantc=True):
    h, w = left.shape
    d16, d32, cf = _outputs(h, w, want16, want32, wantc)
    s = params_struct(p)
    l = np.ascontiguousarray(left, np.uint8)
    r = np.ascontiguousarray(right, np.uint8)
    rc = lib().orc_matcher_compute(w, h, C.byref(s), _ptr(l), _ptr(r), _ptr(d16), _ptr(d32),
                                   _ptr(cf))
    return Outputs(d16, d32, cf, rc)
