"""PFM files, MiddEval3 calib.txt, disparity metrics and benchmark reports.

Python mirror of include/pyrsgm/pyrsgm_io.h (the library's host functions) plus
the reference benchmark's CSV and Markdown report layouts
(reference bench/pyrsgm_bench.cpp:558-714).  Arrays are numpy, on the host.
"""
from __future__ import annotations

import ctypes as C
import os
import statistics
from dataclasses import dataclass, field

import numpy as np

from . import lib


class Metrics(C.Structure):
    """pyrsgm_metrics_c (reference bench/middlebury.h:66-78)."""
    _fields_ = [("evaluated", C.c_int), ("valid", C.c_int),
                ("avg_err", C.c_double), ("rms_err", C.c_double),
                ("bad05", C.c_double), ("bad1", C.c_double), ("bad2", C.c_double),
                ("bad4", C.c_double), ("density", C.c_double), ("invalid", C.c_double),
                ("bad1_all", C.c_double), ("bad2_all", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_bound = False


def _L():
    global _bound
    L = lib()
    if not _bound:
        vp, sz, ci = C.c_void_p, C.c_size_t, C.c_int
        L.pyrsgm_read_pfm.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_float)),
                                      C.POINTER(ci), C.POINTER(ci), C.c_char_p, ci]
        L.pyrsgm_write_pfm.argtypes = [C.c_char_p, vp, sz, ci, ci]
        L.pyrsgm_free.argtypes = [vp]
        L.pyrsgm_free.restype = None
        L.pyrsgm_read_ndisp.argtypes = [C.c_char_p]
        L.pyrsgm_disp16_to_float.argtypes = [vp, sz, vp, sz, ci, ci]
        L.pyrsgm_disp16_to_float.restype = None
        L.pyrsgm_evaluate.argtypes = [vp, sz, vp, sz, vp, sz, ci, ci, C.c_float,
                                      C.POINTER(Metrics)]
        _bound = True
    return L


def _enc(path) -> bytes:
    return os.fsencode(path)


def read_pfm(path) -> np.ndarray:
    """Grayscale PFM as a float32 [H, W] array, top row first (inf = unknown)."""
    data, w, h = C.POINTER(C.c_float)(), C.c_int(), C.c_int()
    err = C.create_string_buffer(512)
    if _L().pyrsgm_read_pfm(_enc(path), C.byref(data), C.byref(w), C.byref(h), err, 512) != 0:
        raise ValueError(err.value.decode())
    try:
        return np.ctypeslib.as_array(data, shape=(h.value, w.value)).copy()
    finally:
        _L().pyrsgm_free(data)


def write_pfm(path, disp: np.ndarray) -> None:
    """Little-endian grayscale PFM (reference examples/stereo_pair.cpp:38-45)."""
    a = np.ascontiguousarray(disp, dtype=np.float32)
    if a.ndim != 2:
        raise ValueError("write_pfm: expected a 2-D array")
    if _L().pyrsgm_write_pfm(_enc(path), a.ctypes.data, a.strides[0], a.shape[1], a.shape[0]) != 0:
        raise OSError(f"write_pfm: cannot write {path}")


def read_ndisp(path) -> int:
    """ndisp= of a MiddEval3 calib.txt (0 if absent)."""
    return _L().pyrsgm_read_ndisp(_enc(path))


def disp16_to_float(disp16: np.ndarray) -> np.ndarray:
    """Fixed-point disparity*16 to pixels; values <= 0 become 0 (invalid)."""
    a = np.ascontiguousarray(disp16, dtype=np.int16)
    out = np.empty(a.shape, np.float32)
    _L().pyrsgm_disp16_to_float(a.ctypes.data, a.strides[0], out.ctypes.data, out.strides[0],
                                a.shape[1], a.shape[0])
    return out


def evaluate(disp: np.ndarray, gt: np.ndarray, mask: np.ndarray | None = None,
             max_disp: float = 1e30) -> dict:
    """Middlebury metrics of a float disparity map against ground truth
    (reference bench/middlebury.h:80-120)."""
    d = np.ascontiguousarray(disp, dtype=np.float32)
    g = np.ascontiguousarray(gt, dtype=np.float32)
    if d.shape != g.shape or d.ndim != 2:
        raise ValueError("evaluate: disp and gt must be 2-D arrays of one shape")
    mk = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    if mk is not None and mk.shape != g.shape:
        raise ValueError("evaluate: mask shape differs")
    m = Metrics()
    _L().pyrsgm_evaluate(d.ctypes.data, d.strides[0], g.ctypes.data, g.strides[0],
                         None if mk is None else mk.ctypes.data, 0 if mk is None else mk.strides[0],
                         g.shape[1], g.shape[0], float(max_disp), C.byref(m))
    return m.as_dict()


# ----------------------------------------------------------- bench reports
@dataclass
class Result:
    """One (scene, algorithm) row of a benchmark (reference bench/pyrsgm_bench.cpp:423-441)."""
    scene: str
    algo: str
    width: int
    height: int
    D: int
    gpu_ms: list = field(default_factory=list)
    wall_ms: list = field(default_factory=list)
    metrics: dict | None = None
    error: str = ""

    @property
    def failed(self) -> bool:
        return bool(self.error)

    def summary(self):
        if not self.gpu_ms:
            return 0.0, 0.0, 0.0, 0.0
        return (statistics.median(self.gpu_ms), sum(self.gpu_ms) / len(self.gpu_ms),
                min(self.gpu_ms), statistics.median(self.wall_ms) if self.wall_ms else 0.0)


def _f(v, prec=2):
    return f"{v:.{prec}f}"


CSV_HEADER = ("scene,algo,width,height,D,gpu_ms_median,gpu_ms_mean,gpu_ms_min,wall_ms_median,"
              "density_pct,bad0.5_pct,bad1_pct,bad2_pct,bad4_pct,avg_err_px,rms_err_px,"
              "bad1_all_pct,bad2_all_pct,error")


def csv_rows(results) -> str:
    """The reference benchmark's CSV (bench/pyrsgm_bench.cpp:558-576)."""
    lines = [CSV_HEADER]
    for r in results:
        med, mean, mn, wall = r.summary()
        row = [r.scene, r.algo, str(r.width), str(r.height), str(r.D), _f(med, 3), _f(mean, 3),
               _f(mn, 3), _f(wall, 3)]
        m = r.metrics
        if m:
            row += [_f(m["density"]), _f(m["bad05"]), _f(m["bad1"]), _f(m["bad2"]), _f(m["bad4"]),
                    _f(m["avg_err"], 3), _f(m["rms_err"], 3), _f(m["bad1_all"]), _f(m["bad2_all"])]
        else:
            row += [""] * 9
        row.append(r.error)
        lines.append(",".join(row))
    return "\n".join(lines) + "\n"


def markdown_summary(results, descriptions: dict | None = None) -> str:
    """The reference benchmark's per-algorithm summary table (bench/pyrsgm_bench.cpp:586-672):
    per-scene medians averaged over an algorithm's successful scenes, FPS = 1000 / that
    average, metrics summed over the scenes that have them and divided by all its scenes,
    paths read from the description ("8 paths" / "4 paths", else 3)."""
    descriptions = descriptions or {}
    algos = []
    for r in results:
        if r.algo not in algos:
            algos.append(r.algo)
    any_metrics = any(r.metrics for r in results if not r.failed)
    rows = []
    for a in algos:
        rs = [r for r in results if r.algo == a and not r.failed]
        n = len(rs)
        if not n:
            continue
        gpu = sum(r.summary()[0] for r in rs) / n
        wall = sum(r.summary()[3] for r in rs) / n
        ms = [r.metrics for r in rs if r.metrics]
        desc = descriptions.get(a, "")
        paths = "8" if "8 paths" in desc else ("4" if "4 paths" in desc else "3")
        row = f"| {a} | {paths} | {_f(gpu)} | {_f(1000.0 / gpu if gpu > 0 else 0.0, 1)} | {_f(wall)}"
        if any_metrics:
            if ms:
                avg = {k: sum(m[k] for m in ms) / n for k in
                       ("density", "bad1", "bad2", "avg_err", "bad1_all", "bad2_all")}
                row += (f" | {_f(avg['density'], 1)} | {_f(avg['bad1'])} | {_f(avg['bad2'])} | "
                        f"{_f(avg['avg_err'])} | {_f(avg['bad1_all'])} | {_f(avg['bad2_all'])}")
            else:
                row += " | - | - | - | - | - | -"
        rows.append(row + " |")
    head = "| algorithm | paths | avg GPU ms | FPS | wall ms"
    sep = "|---|---|---:|---:|---:"
    if any_metrics:
        head += (" | density % | bad1 % (valid) | bad2 % (valid) | avg err px | bad1 % (all) |"
                 " bad2 % (all)")
        sep += "|---:|---:|---:|---:|---:|---:"
    out = [head + " |", sep + "|", *rows, ""]
    if descriptions:
        out.append("### Algorithm settings\n")
        out += [f"- `{a}`: {descriptions[a]}" for a in algos if a in descriptions]
        out.append("")
    return "\n".join(out) + "\n"
