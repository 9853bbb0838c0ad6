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


class _ServeRequestC(C.Structure):
    _fields_ = [("left", C.c_void_p), ("left_pitch", C.c_size_t),
                ("right", C.c_void_p), ("right_pitch", C.c_size_t),
                ("width", C.c_int), ("height", C.c_int),
                ("disp16", C.c_void_p), ("disp16_pitch", C.c_size_t),
                ("disp32", C.c_void_p), ("disp32_pitch", C.c_size_t),
                ("conf", C.c_void_p), ("conf_pitch", C.c_size_t)]


class ServeResult(C.Structure):
    """pyrsgm_serve_result_c: a finished request (times in steady-clock ns)."""
    _fields_ = [("ticket", C.c_ulonglong), ("status", C.c_int), ("batch_size", C.c_int),
                ("arrival_ns", C.c_longlong), ("launch_ns", C.c_longlong),
                ("done_ns", C.c_longlong)]

    @property
    def latency_ms(self) -> float:
        return (self.done_ns - self.arrival_ns) * 1e-6


class ServeStats(C.Structure):
    """pyrsgm_serve_stats_c."""
    _fields_ = [("submitted", C.c_ulonglong), ("completed", C.c_ulonglong),
                ("failed", C.c_ulonglong), ("pairs_per_s", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("max_ms", C.c_double), ("mean_ms", C.c_double),
                ("mean_batch", C.c_double), ("resident_bytes", C.c_size_t),
                ("peak_bytes", C.c_size_t), ("budget_bytes", C.c_size_t), ("lanes", C.c_int),
                ("lanes_created", C.c_int), ("lanes_evicted", C.c_int), ("launches", C.c_int)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


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
        L.pyrsgm_compute_single_scale_noalloc.argtypes = [vp] * 5 + [sz] * 5 + [ci] * 14 + [cf] * 4 + [vp]
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
        L.pyrsgm_batch_compute.argtypes = [vp] + [vp, sz, sz] * 5 + [ci, vp, C.c_char_p, ci]
        L.pyrsgm_batch_capture.argtypes = [vp] + [vp, sz, sz] * 5 + [ci, C.c_char_p, ci]
        L.pyrsgm_batch_required_bytes.argtypes = [ci, ci, ci, C.POINTER(_ParamsC), C.c_char_p, ci]
        L.pyrsgm_batch_required_bytes.restype = sz
        L.pyrsgm_batch_launch.argtypes = [vp, vp, C.c_char_p, ci]
        L.pyrsgm_batch_set_tuning.argtypes = [vp, C.POINTER(ci), C.c_char_p, ci]
        L.pyrsgm_batch_get_tuning.argtypes = [vp, C.POINTER(ci)]
        L.pyrsgm_batch_tune.argtypes = ([vp] + [vp, sz, sz] * 5 +
                                        [vp, ci, C.POINTER(ci), C.POINTER(C.c_double), C.c_char_p, ci])
        L.pyrsgm_batch_device_bytes.argtypes = [vp]
        L.pyrsgm_batch_device_bytes.restype = sz
        ll, ull = C.c_longlong, C.c_ulonglong
        L.pyrsgm_serve_create.argtypes = [sz, ci, ci, ci, C.c_char_p, ci]
        L.pyrsgm_serve_create.restype = vp
        L.pyrsgm_serve_destroy.argtypes = [vp]
        L.pyrsgm_serve_destroy.restype = None
        L.pyrsgm_serve_submit.argtypes = [vp, C.POINTER(_ServeRequestC), C.POINTER(_ParamsC), ll,
                                          C.c_char_p, ci]
        L.pyrsgm_serve_submit.restype = ll
        L.pyrsgm_serve_poll.argtypes = [vp, ull, C.POINTER(ServeResult)]
        L.pyrsgm_serve_wait.argtypes = [vp, ull, C.POINTER(ServeResult), C.c_char_p, ci]
        L.pyrsgm_serve_drain.argtypes = [vp]
        L.pyrsgm_serve_take.argtypes = [vp, C.POINTER(ServeResult), ci]
        L.pyrsgm_serve_stats.argtypes = [vp, C.POINTER(ServeStats)]
        L.pyrsgm_serve_reset_stats.argtypes = [vp]
        L.pyrsgm_serve_reset_stats.restype = None
        L.pyrsgm_serve_now_ns.restype = ll
        L.pyrsgm_mixed_create.argtypes = [ci, C.POINTER(ci), C.POINTER(ci), C.POINTER(_ParamsC),
                                          C.c_char_p, ci]
        L.pyrsgm_mixed_create.restype = vp
        L.pyrsgm_mixed_destroy.argtypes = [vp]
        L.pyrsgm_mixed_destroy.restype = None
        L.pyrsgm_mixed_compute.argtypes = [vp, C.POINTER(_PairIOC), vp, C.c_char_p, ci]
        L.pyrsgm_mixed_capture.argtypes = [vp, C.POINTER(_PairIOC), C.c_char_p, ci]
        L.pyrsgm_mixed_launch.argtypes = [vp, vp, C.c_char_p, ci]
        L.pyrsgm_mixed_set_tuning.argtypes = [vp, C.POINTER(MixedTuning), C.c_char_p, ci]
        L.pyrsgm_mixed_get_tuning.argtypes = [vp, C.POINTER(MixedTuning)]
        L.pyrsgm_mixed_tune.argtypes = [vp, C.POINTER(_PairIOC), vp, ci, C.POINTER(MixedTuning),
                                        C.POINTER(C.c_double), C.c_char_p, ci]
        L.pyrsgm_mixed_plan.argtypes = [vp, C.POINTER(ll)]
        L.pyrsgm_mixed_device_bytes.argtypes = [vp]
        L.pyrsgm_mixed_device_bytes.restype = sz
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


def compute_single_scale_noalloc(d_left_gray, d_right_gray, d_disp, d_disp_float, d_conf_float,
                                 left_step_bytes, right_step_bytes, disp_step_bytes,
                                 disp_float_step_bytes, conf_float_step_bytes, width, height,
                                 numDisparities, minDisparity, blockSize, P1, P2,
                                 uniquenessRatio, disp12MaxDiff, preFilterCap, costMode,
                                 postFlags, numPaths, adaptiveP2, p2Alpha, p2Beta, p2LowerMult,
                                 p2UpperMult, stream=0) -> int:
    """pyrsgm_compute_single_scale on the library's fixed device scratch array
    (no allocation; NOALLOC_SCRATCH_BYTES per device; overlapping calls on one
    device race)."""
    return lib().pyrsgm_compute_single_scale_noalloc(
        d_left_gray, d_right_gray, d_disp, d_disp_float, d_conf_float, left_step_bytes,
        right_step_bytes, disp_step_bytes, disp_float_step_bytes, conf_float_step_bytes, width,
        height, numDisparities, minDisparity, blockSize, P1, P2, uniquenessRatio, disp12MaxDiff,
        preFilterCap, costMode, postFlags, numPaths, adaptiveP2, p2Alpha, p2Beta, p2LowerMult,
        p2UpperMult, stream)


NOALLOC_SCRATCH_BYTES = 192 << 20  # PYRSGM_NOALLOC_SCRATCH_BYTES


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


def _check(t, name, dtype, shape):
    """Raises ValueError unless `t` is a device tensor of `dtype` whose shape is
    `shape` (None entries match any size) with contiguous rows (last stride 1)
    and a non-negative row pitch.  The kernels address rows by pointer and
    pitch only, so a smaller, strided or differently typed tensor would be
    read or written out of bounds."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"pyrsgm: {name} must be a torch tensor, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"pyrsgm: {name} must be a device tensor")
    if t.dtype != dtype:
        raise ValueError(f"pyrsgm: {name} must be {dtype}, got {t.dtype}")
    if t.dim() != len(shape) or any(s is not None and s != n for s, n in zip(shape, t.shape)):
        want = ", ".join("N" if s is None else str(s) for s in shape)
        raise ValueError(f"pyrsgm: {name} must have shape [{want}], got {list(t.shape)}")
    if t.stride(-1) != 1 or any(t.stride(i) < 0 for i in range(t.dim())):
        raise ValueError(f"pyrsgm: {name} rows must be contiguous (last stride 1), "
                         f"got strides {t.stride()}")
    for i in range(t.dim() - 1):
        # Rows (and pairs) must not overlap: each outer stride spans the inner extent.
        if t.shape[i] > 1 and t.stride(i) < t.shape[i + 1] * t.stride(i + 1):
            raise ValueError(f"pyrsgm: {name} has overlapping rows or pairs (strides {t.stride()}, "
                             f"shape {list(t.shape)})")


# Tensor roles and their element types.
_ROLE_DTYPES = {"left": "uint8", "right": "uint8", "disp16": "int16", "disp32": "float32",
                "conf": "float32"}


def _checked(role, t, shape):
    if t is not None:
        import torch
        _check(t, role, getattr(torch, _ROLE_DTYPES[role]), shape)
    return t


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
        hw = (self._hgt, self._w)
        for role, t in (("left", left), ("right", right)):
            if t is None:
                raise ValueError(f"pyrsgm: {role} image is required")
            _checked(role, t, hw)
        for role, t in (("disp16", disp16), ("disp32", disp32), ("conf", conf)):
            _checked(role, t, hw)
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

    @staticmethod
    def required_bytes(width: int, height: int, batch: int, params: Params) -> int:
        """Device memory a BatchMatcher of these arguments allocates."""
        err = C.create_string_buffer(512)
        n = lib().pyrsgm_batch_required_bytes(width, height, batch, C.byref(_to_c(params)), err, 512)
        if n == 0:
            raise ValueError(err.value.decode())
        return n

    def _args(self, left, right, disp16, disp32, conf, count):
        count = self.batch if count is None else count
        if not 1 <= count <= self.batch:
            raise ValueError(f"pyrsgm: pair count {count} outside [1, {self.batch}]")
        for role, t in (("left", left), ("right", right), ("disp16", disp16),
                        ("disp32", disp32), ("conf", conf)):
            if t is None:
                if role in ("left", "right"):
                    raise ValueError(f"pyrsgm: {role} images are required")
                continue
            _checked(role, t, (None, self.height, self.width))
            if t.shape[0] < count or t.shape[0] > self.batch:
                raise ValueError(f"pyrsgm: {role} holds {t.shape[0]} pairs; expected "
                                 f"{count}..{self.batch}")
        out = []
        for t in (left, right, disp16, disp32, conf):
            out += list(_bptr(t))
        return out + [count]

    def compute(self, left, right, disp16=None, disp32=None, conf=None, stream=None, count=None):
        """The first `count` pairs (default: all) of [N, H, W] tensors."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        err = C.create_string_buffer(512)
        if lib().pyrsgm_batch_compute(self._h, *self._args(left, right, disp16, disp32, conf, count),
                                      s.cuda_stream, err, 512) != 0:
            raise RuntimeError(err.value.decode())

    def capture(self, left, right, disp16=None, disp32=None, conf=None, count=None):
        """Record compute() on these tensors into a HIP graph; replay with launch()."""
        err = C.create_string_buffer(512)
        if lib().pyrsgm_batch_capture(self._h, *self._args(left, right, disp16, disp32, conf, count),
                                      err, 512) != 0:
            raise RuntimeError(err.value.decode())

    def launch(self, stream=None):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        err = C.create_string_buffer(512)
        if lib().pyrsgm_batch_launch(self._h, s.cuda_stream, err, 512) != 0:
            raise RuntimeError(err.value.decode())

    _TUNING = ("census_rows", "paired", "xcd_remap", "side_stream", "fused", "modular")

    def set_tuning(self, census_rows=4, paired=-1, xcd_remap=-1, side_stream=1, fused=-1,
                   modular=-1):
        """Launch shapes (every choice gives the same bits); drops a captured graph."""
        t = (C.c_int * 6)(census_rows, paired, xcd_remap, side_stream, fused, modular)
        err = C.create_string_buffer(512)
        if lib().pyrsgm_batch_set_tuning(self._h, t, err, 512) != 0:
            raise ValueError(err.value.decode())

    def tuning(self) -> dict:
        t = (C.c_int * 6)()
        lib().pyrsgm_batch_get_tuning(self._h, t)
        return dict(zip(self._TUNING, list(t)))

    def tune(self, left, right, disp16=None, disp32=None, conf=None, reps=3, stream=None):
        """Time candidate launch shapes on all pairs; keeps the fastest.
        Returns (tuning, ms per batch)."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        args = self._args(left, right, disp16, disp32, conf, None)[:-1]
        best, ms = (C.c_int * 6)(), C.c_double()
        err = C.create_string_buffer(512)
        if lib().pyrsgm_batch_tune(self._h, *args, s.cuda_stream, reps, best, C.byref(ms),
                                   err, 512) != 0:
            raise RuntimeError(err.value.decode())
        return dict(zip(self._TUNING, list(best))), ms.value

    def device_memory_bytes(self) -> int:
        return lib().pyrsgm_batch_device_bytes(self._h)


class _PairIOC(C.Structure):
    """pyrsgm_pair_io_c (include/pyrsgm/pyrsgm_mixed.h)."""
    _fields_ = [("left", C.c_void_p), ("left_pitch", C.c_size_t),
                ("right", C.c_void_p), ("right_pitch", C.c_size_t),
                ("disp16", C.c_void_p), ("disp16_pitch", C.c_size_t),
                ("disp32", C.c_void_p), ("disp32_pitch", C.c_size_t),
                ("conf", C.c_void_p), ("conf_pitch", C.c_size_t)]


class MixedTuning(C.Structure):
    """pyrsgm_mixed_tuning_c: streams for the per-pair stages (1..4), resident
    workgroups per CU of the queued path launches (0: occupancy), paired lines
    (-1 by launch size, 0 off, 1 on), and whether pyrDown, descriptors, costs and
    post run as queued launches over all pairs (1) or per pair (0)."""
    _fields_ = [("branches", C.c_int), ("blocks_per_cu", C.c_int), ("paired", C.c_int),
                ("queued_stages", C.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class MixedBatch:
    """pyrsgm::MixedBatch: pairs of different sizes and parameters in one fused
    pass (include/pyrsgm/pyrsgm_mixed.h).

    ``specs`` is a list of (width, height, Params).  Each pair's tensors are
    2-D device tensors (uint8 in; int16 / float32 out)."""

    def __init__(self, specs):
        err = C.create_string_buffer(512)
        n = len(specs)
        ws = (C.c_int * n)(*[int(w) for w, _, _ in specs])
        hs = (C.c_int * n)(*[int(h) for _, h, _ in specs])
        ps = (_ParamsC * n)(*[_to_c(p) for _, _, p in specs])
        h = lib().pyrsgm_mixed_create(n, ws, hs, ps, err, 512)
        if not h:
            msg = err.value.decode()
            raise (ValueError if "pyrsgm:" in msg and "HIP" not in msg else RuntimeError)(msg)
        self._h = h
        self.specs = list(specs)

    def close(self):
        if getattr(self, "_h", None):
            lib().pyrsgm_mixed_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _io(self, pairs):
        """pairs: one (left, right, disp16, disp32, conf) tuple of tensors / None per pair."""
        if len(pairs) != len(self.specs):
            raise ValueError(f"pyrsgm: expected {len(self.specs)} pairs, got {len(pairs)}")
        arr = (_PairIOC * len(pairs))()
        for i, ((w, h, _), tensors) in enumerate(zip(self.specs, pairs)):
            vals = []
            if len(tensors) != 5:
                raise ValueError(f"pyrsgm: pair {i}: expected (left, right, disp16, disp32, conf)")
            for role, t in zip(("left", "right", "disp16", "disp32", "conf"), tensors):
                if t is None and role in ("left", "right"):
                    raise ValueError(f"pyrsgm: pair {i}: {role} image is required")
                try:
                    _checked(role, t, (h, w))
                except ValueError as e:
                    raise ValueError(f"pyrsgm: pair {i}: {e}") from None
                vals += list(_ptr(t))
            for (name, _), v in zip(_PairIOC._fields_, vals):
                setattr(arr[i], name, v)
        return arr

    def compute(self, pairs, stream=None):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        err = C.create_string_buffer(512)
        if lib().pyrsgm_mixed_compute(self._h, self._io(pairs), s.cuda_stream, err, 512) != 0:
            raise RuntimeError(err.value.decode())

    def capture(self, pairs):
        """Record compute() on these tensors into a HIP graph; replay with launch()."""
        err = C.create_string_buffer(512)
        if lib().pyrsgm_mixed_capture(self._h, self._io(pairs), err, 512) != 0:
            raise RuntimeError(err.value.decode())

    def launch(self, stream=None):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        err = C.create_string_buffer(512)
        if lib().pyrsgm_mixed_launch(self._h, s.cuda_stream, err, 512) != 0:
            raise RuntimeError(err.value.decode())

    def set_tuning(self, branches=4, blocks_per_cu=0, paired=-1, queued_stages=1):
        err = C.create_string_buffer(512)
        t = MixedTuning(branches, blocks_per_cu, paired, queued_stages)
        if lib().pyrsgm_mixed_set_tuning(self._h, C.byref(t), err, 512) != 0:
            raise ValueError(err.value.decode())

    def tuning(self) -> dict:
        t = MixedTuning()
        lib().pyrsgm_mixed_get_tuning(self._h, C.byref(t))
        return t.as_dict()

    def tune(self, pairs, reps=3, stream=None):
        """Time the candidate launch shapes; keeps the fastest.  Returns (tuning, ms per pass)."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        err = C.create_string_buffer(512)
        best, ms = MixedTuning(), C.c_double()
        if lib().pyrsgm_mixed_tune(self._h, self._io(pairs), s.cuda_stream, reps, C.byref(best),
                                   C.byref(ms), err, 512) != 0:
            raise RuntimeError(err.value.decode())
        return best.as_dict(), ms.value

    def plan(self) -> dict:
        out = (C.c_longlong * 6)()
        lib().pyrsgm_mixed_plan(self._h, out)
        return {"queued_path_launches": out[0], "path_tasks": out[1], "work_items": out[2],
                "per_pair_stage_calls": out[3], "queued_stage_launches": out[4],
                "stage_tasks": out[5]}

    def device_memory_bytes(self) -> int:
        return lib().pyrsgm_mixed_device_bytes(self._h)


class Server:
    """pyrsgm::Server: the serving loop (include/pyrsgm/pyrsgm_serve.h).

    Requests are 2-D device tensors (uint8 in; int16 / float32 out) that must
    stay alive and unchanged until the request is finished."""

    def __init__(self, memory_budget: int = 0, max_batch: int = 16, lanes_per_key: int = 2,
                 use_graphs: bool = True):
        err = C.create_string_buffer(512)
        h = lib().pyrsgm_serve_create(memory_budget, max_batch, lanes_per_key, int(use_graphs),
                                      err, 512)
        if not h:
            msg = err.value.decode()
            raise (ValueError if "must be" in msg else RuntimeError)(msg)
        self._h = h
        self._keep = {}  # ticket -> tensors the device work reads or writes

    def close(self):
        if getattr(self, "_h", None):
            lib().pyrsgm_serve_destroy(self._h)  # finishes everything submitted
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def now_ns() -> int:
        return lib().pyrsgm_serve_now_ns()

    def submit(self, left, right, params: Params, disp16=None, disp32=None, conf=None,
               arrival_ns: int = 0) -> int:
        """Queue one pair; returns its ticket.  arrival_ns (steady clock) defaults to now."""
        if left is None or right is None:
            raise ValueError("pyrsgm: left and right images are required")
        _checked("left", left, (None, None))
        h, w = left.shape
        req = _ServeRequestC()
        req.width, req.height = w, h
        for name, t in (("left", left), ("right", right), ("disp16", disp16),
                        ("disp32", disp32), ("conf", conf)):
            _checked(name, t, (h, w))
            ptr, pitch = _ptr(t)
            setattr(req, name, ptr)
            setattr(req, name + "_pitch", pitch)
        err = C.create_string_buffer(512)
        t = lib().pyrsgm_serve_submit(self._h, C.byref(req), C.byref(_to_c(params)), arrival_ns,
                                      err, 512)
        if t < 0:
            raise ValueError(err.value.decode())
        self._keep[t] = (left, right, disp16, disp32, conf)
        return t

    def poll(self, ticket: int):
        """The ServeResult once the ticket is finished, else None."""
        out = ServeResult()
        if lib().pyrsgm_serve_poll(self._h, ticket, C.byref(out)) == 1:
            self._keep.pop(ticket, None)
            return out
        return None

    def wait(self, ticket: int) -> ServeResult:
        out = ServeResult()
        err = C.create_string_buffer(512)
        if lib().pyrsgm_serve_wait(self._h, ticket, C.byref(out), err, 512) != 0:
            raise ValueError(err.value.decode())
        self._keep.pop(ticket, None)
        return out

    def drain(self):
        lib().pyrsgm_serve_drain(self._h)

    def take(self, max_results: int = 4096) -> list:
        """Finished results not yet polled (at most max_results)."""
        arr = (ServeResult * max_results)()
        n = lib().pyrsgm_serve_take(self._h, arr, max_results)
        out = [arr[i] for i in range(max(n, 0))]
        for r in out:
            self._keep.pop(r.ticket, None)
        return out

    def stats(self) -> ServeStats:
        out = ServeStats()
        lib().pyrsgm_serve_stats(self._h, C.byref(out))
        return out

    def reset_stats(self):
        lib().pyrsgm_serve_reset_stats(self._h)

