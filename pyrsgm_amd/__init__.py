"""pyrsgm_amd -- MI355X (gfx950) coarse-to-fine SGM stereo engine.

Python host mirror of gegeges/pyrsgm's two interfaces, both backed by the HIP
library ``pyrsgm_amd/lib/libpyrsgm.so`` (build with ``python -m
pyrsgm_amd.build``):

* the C-ABI of ``pyrsgm_c.h`` -- :func:`create_workspace`,
  :func:`workspace_reserve`, :func:`compute_single_scale`,
  :func:`compute_level`, :func:`pyrdown_u8` -- same argument names, order and
  meaning, device pointers as integers, pitches in bytes, return 0 / -1;
* the C++ ``pyrsgm::StereoMatcher`` -- :class:`Params`, :func:`validate_params`,
  :class:`StereoMatcher` with ``compute`` / ``level_info`` / ``num_levels`` /
  ``device_memory_bytes``; errors raise ``ValueError`` (the C++
  std::invalid_argument) or ``RuntimeError``.

There is no CPU fallback: every call goes to the HIP library and fails loudly
if it is missing.  torch is used only to hand device pointers and streams in.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from dataclasses import dataclass, fields

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PYRSGM_LIB", os.path.join(HERE, "lib", "libpyrsgm.so"))

PYRSGM_COST_BT = 0
PYRSGM_COST_CENSUS = 1
PYRSGM_POST_LRCHECK = 1
PYRSGM_POST_MEDIAN = 2

_lib = None


class CostMode(enum.IntEnum):
    BirchfieldTomasi = 0
    Census = 1


@dataclass
class Params:
    """Field-for-field mirror of pyrsgm::Params (reference pyrsgm.h:32-82)."""
    num_disparities: int = 128
    min_disparity: int = 0
    block_size: int = 5
    cost_mode: CostMode = CostMode.BirchfieldTomasi
    p1: int = 300
    p2: int = 2400
    uniqueness_ratio: int = 10
    disp12_max_diff: int = 1
    lr_check: bool = True
    median_filter: bool = True
    num_paths: int = 3
    pre_filter_cap: int = 63
    pyramid_levels: int = 0
    pyramid_coarse_disparities: int = 128
    pyramid_max_disparity: int = 0
    pyramid_min_confidence: float = 0.05
    adaptive_p2: bool = False
    p2_alpha: float = 3.041
    p2_beta: float = 16.137
    p2_lower_mult: float = 0.601
    p2_upper_mult: float = 1.77


@dataclass
class LevelInfo:
    level: int
    width: int
    height: int
    num_labels: int
    max_disparity: int
    block_size: int
    refine: bool


class _ParamsC(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "num_disparities", "min_disparity", "block_size", "cost_mode", "p1", "p2",
        "uniqueness_ratio", "disp12_max_diff", "lr_check", "median_filter", "num_paths",
        "pre_filter_cap", "pyramid_levels", "pyramid_coarse_disparities",
        "pyramid_max_disparity")] + [
        ("pyramid_min_confidence", C.c_float), ("adaptive_p2", C.c_int),
        ("p2_alpha", C.c_float), ("p2_beta", C.c_float), ("p2_lower_mult", C.c_float),
        ("p2_upper_mult", C.c_float)]


def _to_c(p: Params) -> _ParamsC:
    s = _ParamsC()
    for f in fields(Params):
        v = getattr(p, f.name)
        setattr(s, f.name, float(v) if isinstance(getattr(s, f.name), float) else int(v))
    return s


def _from_c(s: _ParamsC) -> Params:
    kw = {}
    for f in fields(Params):
        v = getattr(s, f.name)
        if f.name == "cost_mode":
            v = CostMode(v)
        elif f.type in ("bool", bool):
            v = bool(v)
        kw[f.name] = v
    return Params(**kw)


def lib() -> C.CDLL:
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"pyrsgm_amd: native library missing ({LIB_PATH}); "
                               "build it with `python -m pyrsgm_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, sz, ci, cf = C.c_void_p, C.c_size_t, C.c_int, C.c_float
        L.pyrsgm_create_workspace.restype = vp
        L.pyrsgm_destroy_workspace.argtypes = [vp]
        L.pyrsgm_destroy_workspace.restype = None
        L.pyrsgm_workspace_reserve.argtypes = [vp, ci, ci, ci]
        L.pyrsgm_compute_single_scale.argtypes = [vp] * 6 + [sz] * 5 + [ci] * 14 + [cf] * 4 + [vp]
        L.pyrsgm_compute_level.argtypes = [vp] * 8 + [sz] * 7 + [ci] * 17 + [cf, ci] + [cf] * 4 + [vp]
        L.pyrsgm_pyrdown_u8.argtypes = [vp, sz, ci, ci, vp, sz, ci, ci, vp]
        L.pyrsgm_validate_params.argtypes = [C.POINTER(_ParamsC), C.POINTER(_ParamsC), C.c_char_p, ci]
        L.pyrsgm_matcher_create.argtypes = [ci, ci, C.POINTER(_ParamsC), C.c_char_p, ci]
        L.pyrsgm_matcher_create.restype = vp
        L.pyrsgm_matcher_destroy.argtypes = [vp]
        L.pyrsgm_matcher_destroy.restype = None
        L.pyrsgm_matcher_compute.argtypes = [vp, vp, sz, vp, sz, vp, sz, vp, sz, vp, sz, vp,
                                             C.c_char_p, ci]
        L.pyrsgm_matcher_level_info.argtypes = [vp, ci, C.POINTER(C.c_int)]
        L.pyrsgm_matcher_num_levels.argtypes = [vp]
        L.pyrsgm_matcher_params.argtypes = [vp, C.POINTER(_ParamsC)]
        L.pyrsgm_matcher_device_bytes.argtypes = [vp]
        L.pyrsgm_matcher_device_bytes.restype = sz
        L.pyrsgm_version.restype = C.c_char_p
        L.pyrsgm_batch_create.argtypes = [ci, ci, ci, C.POINTER(_ParamsC), C.c_char_p, ci]
        L.pyrsgm_batch_create.restype = vp
        L.pyrsgm_batch_destroy.argtypes = [vp]
        L.pyrsgm_batch_destroy.restype = None
        L.pyrsgm_batch_compute.argtypes = [vp] + [vp, sz, sz] * 5 + [vp, C.c_char_p, ci]
        L.pyrsgm_batch_capture.argtypes = [vp] + [vp, sz, sz] * 5 + [C.c_char_p, ci]
        L.pyrsgm_batch_launch.argtypes = [vp, vp, C.c_char_p, ci]
        L.pyrsgm_batch_device_bytes.argtypes = [vp]
        L.pyrsgm_batch_device_bytes.restype = sz
        _lib = L
    return _lib


def version() -> str:
    return lib().pyrsgm_version().decode()


# ------------------------------------------------------------------ C-ABI
def create_workspace() -> int:
    return lib().pyrsgm_create_workspace()


def destroy_workspace(ws: int) -> None:
    lib().pyrsgm_destroy_workspace(ws)


def workspace_reserve(ws, width, height, num_disparities) -> int:
    return lib().pyrsgm_workspace_reserve(ws, width, height, num_disparities)


def compute_single_scale(workspace, d_left_gray, d_right_gray, d_disp, d_disp_float,
                         d_conf_float, left_step_bytes, right_step_bytes, disp_step_bytes,
                         disp_float_step_bytes, conf_float_step_bytes, width, height,
                         numDisparities, minDisparity, blockSize, P1, P2, uniquenessRatio,
                         disp12MaxDiff, preFilterCap, costMode, postFlags, numPaths, adaptiveP2,
                         p2Alpha, p2Beta, p2LowerMult, p2UpperMult, stream=0) -> int:
    return lib().pyrsgm_compute_single_scale(
        workspace, d_left_gray, d_right_gray, d_disp, d_disp_float, d_conf_float,
        left_step_bytes, right_step_bytes, disp_step_bytes, disp_float_step_bytes,
        conf_float_step_bytes, width, height, numDisparities, minDisparity, blockSize, P1, P2,
        uniquenessRatio, disp12MaxDiff, preFilterCap, costMode, postFlags, numPaths, adaptiveP2,
        p2Alpha, p2Beta, p2LowerMult, p2UpperMult, stream)


def compute_level(workspace, d_left_gray, d_right_gray, d_parent_disp, d_parent_conf, d_disp,
                  d_disp_float, d_conf_float, left_step_bytes, right_step_bytes,
                  parent_disp_step_bytes, parent_conf_step_bytes, disp_step_bytes,
                  disp_float_step_bytes, conf_float_step_bytes, width, height, parent_width,
                  parent_height, numDisparities, minDisparity, maxDisparity, blockSize, P1, P2,
                  uniquenessRatio, disp12MaxDiff, preFilterCap, costMode, postFlags, numPaths,
                  refineLevel, minConfidence, adaptiveP2, p2Alpha, p2Beta, p2LowerMult,
                  p2UpperMult, stream=0) -> int:
    return lib().pyrsgm_compute_level(
        workspace, d_left_gray, d_right_gray, d_parent_disp, d_parent_conf, d_disp, d_disp_float,
        d_conf_float, left_step_bytes, right_step_bytes, parent_disp_step_bytes,
        parent_conf_step_bytes, disp_step_bytes, disp_float_step_bytes, conf_float_step_bytes,
        width, height, parent_width, parent_height, numDisparities, minDisparity, maxDisparity,
        blockSize, P1, P2, uniquenessRatio, disp12MaxDiff, preFilterCap, costMode, postFlags,
        numPaths, refineLevel, minConfidence, adaptiveP2, p2Alpha, p2Beta, p2LowerMult,
        p2UpperMult, stream)


def pyrdown_u8(d_src, src_step_bytes, src_width, src_height, d_dst, dst_step_bytes, dst_width,
               dst_height, stream=0) -> int:
    return lib().pyrsgm_pyrdown_u8(d_src, src_step_bytes, src_width, src_height, d_dst,
                                   dst_step_bytes, dst_width, dst_height, stream)


# ------------------------------------------------------------- C++ matcher
def validate_params(p: Params) -> Params:
    """pyrsgm::validateParams; raises ValueError like std::invalid_argument."""
    out = _ParamsC()
    err = C.create_string_buffer(512)
    if lib().pyrsgm_validate_params(C.byref(_to_c(p)), C.byref(out), err, 512) != 0:
        raise ValueError(err.value.decode())
    return _from_c(out)


def invalid_disparity16(p: Params) -> int:
    v = ((p.min_disparity - 1) * 16) & 0xFFFF
    return v - 0x10000 if v >= 0x8000 else v


def _ptr(t):
    if t is None:
        return None, 0
    return t.data_ptr(), t.stride(0) * t.element_size()


class StereoMatcher:
    """pyrsgm::StereoMatcher for one image size (reference pyrsgm.h:116-144)."""

    def __init__(self, width: int, height: int, params: Params):
        err = C.create_string_buffer(512)
        h = lib().pyrsgm_matcher_create(width, height, C.byref(_to_c(params)), err, 512)
        if not h:
            msg = err.value.decode()
            raise (ValueError if "pyrsgm:" in msg and "HIP" not in msg else RuntimeError)(msg)
        self._h = h
        self._w, self._hgt = width, height

    def close(self):
        if getattr(self, "_h", None):
            lib().pyrsgm_matcher_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def compute_ptr(self, d_left, left_pitch, d_right, right_pitch, d_disp16=None, disp16_pitch=0,
                    d_disp32=None, disp32_pitch=0, d_conf=None, conf_pitch=0, stream=0):
        err = C.create_string_buffer(512)
        rc = lib().pyrsgm_matcher_compute(self._h, d_left, left_pitch, d_right, right_pitch,
                                          d_disp16, disp16_pitch, d_disp32, disp32_pitch, d_conf,
                                          conf_pitch, stream, err, 512)
        if rc != 0:
            msg = err.value.decode()
            raise (ValueError if "null input" in msg else RuntimeError)(msg)

    def compute(self, left, right, disp16=None, disp32=None, conf=None, stream=None):
        """Torch-tensor convenience: 2-D device tensors (uint8 in, int16/float32 out)."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        l, lp = _ptr(left)
        r, rp = _ptr(right)
        d16, p16 = _ptr(disp16)
        d32, p32 = _ptr(disp32)
        dc, pc = _ptr(conf)
        self.compute_ptr(l, lp, r, rp, d16, p16, d32, p32, dc, pc, s.cuda_stream)

    def width(self) -> int:
        return self._w

    def height(self) -> int:
        return self._hgt

    def params(self) -> Params:
        out = _ParamsC()
        lib().pyrsgm_matcher_params(self._h, C.byref(out))
        return _from_c(out)

    def num_levels(self) -> int:
        return lib().pyrsgm_matcher_num_levels(self._h)

    def level_info(self, level: int) -> LevelInfo:
        out = (C.c_int * 7)()
        if lib().pyrsgm_matcher_level_info(self._h, level, out) != 0:
            raise IndexError("pyrsgm: level")
        return LevelInfo(*out[:6], bool(out[6]))

    def device_memory_bytes(self) -> int:
        return lib().pyrsgm_matcher_device_bytes(self._h)


def _bptr(t):
    """(pointer, row pitch, pair stride) of a [N, H, W] device tensor (or None)."""
    if t is None:
        return None, 0, 0
    es = t.element_size()
    return t.data_ptr(), t.stride(1) * es, t.stride(0) * es


class BatchMatcher:
    """pyrsgm::BatchMatcher: N pairs of one size per launch (include/pyrsgm/pyrsgm_batch.h).

    Tensors are [N, H, W] device tensors (uint8 in, int16 / float32 out), each
    pair's rows contiguous up to the row stride."""

    def __init__(self, width: int, height: int, batch: int, params: Params):
        err = C.create_string_buffer(512)
        h = lib().pyrsgm_batch_create(width, height, batch, C.byref(_to_c(params)), err, 512)
        if not h:
            msg = err.value.decode()
            raise (ValueError if "pyrsgm:" in msg and "HIP" not in msg else RuntimeError)(msg)
        self._h = h
        self.width, self.height, self.batch = width, height, batch

    def close(self):
        if getattr(self, "_h", None):
            lib().pyrsgm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _args(self, left, right, disp16, disp32, conf):
        for t in (left, right, disp16, disp32, conf):
            if t is not None and (t.dim() != 3 or t.shape[0] != self.batch or
                                  t.shape[1] != self.height or t.shape[2] != self.width):
                raise ValueError(f"pyrsgm: expected [{self.batch}, {self.height}, {self.width}] "
                                 f"tensors, got {tuple(t.shape)}")
        out = []
        for t in (left, right, disp16, disp32, conf):
            out += list(_bptr(t))
        return out

    def compute(self, left, right, disp16=None, disp32=None, conf=None, stream=None):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        err = C.create_string_buffer(512)
        if lib().pyrsgm_batch_compute(self._h, *self._args(left, right, disp16, disp32, conf),
                                      s.cuda_stream, err, 512) != 0:
            raise RuntimeError(err.value.decode())

    def capture(self, left, right, disp16=None, disp32=None, conf=None):
        """Record compute() on these tensors into a HIP graph; replay with launch()."""
        err = C.create_string_buffer(512)
        if lib().pyrsgm_batch_capture(self._h, *self._args(left, right, disp16, disp32, conf),
                                      err, 512) != 0:
            raise RuntimeError(err.value.decode())

    def launch(self, stream=None):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        err = C.create_string_buffer(512)
        if lib().pyrsgm_batch_launch(self._h, s.cuda_stream, err, 512) != 0:
            raise RuntimeError(err.value.decode())

    def device_memory_bytes(self) -> int:
        return lib().pyrsgm_batch_device_bytes(self._h)
