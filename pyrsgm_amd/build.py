"""Build recipe of the native library: hipcc for gfx950, in-tree.

Produces ``pyrsgm_amd/lib/libpyrsgm.so`` from ``pyrsgm_amd/csrc``.  Objects go
to ``pyrsgm_amd/lib/obj`` (both git-ignored).  ``python -m pyrsgm_amd.build``
runs it; ``__graft_entry__.build()`` calls :func:`build`.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.environ.get("PYRSGM_LIBDIR", os.path.join(HERE, "lib"))
OBJDIR = os.path.join(LIBDIR, "obj")
LIB = os.path.join(LIBDIR, "libpyrsgm.so")

# kernels_path.hip once per (mode, adaptive P2, volume form, per-launch |
# queued): about 170 kernel instantiations in 32 units that compile in
# parallel.  Single scale has no exact form.
PATH_PARTS = [
    ("kernels_path.hip", f"kernels_path_m{m}a{a}f{f}{'q' if q else ''}",
     {"PYRSGM_PATH_MODE": m, "PYRSGM_PATH_ADAPT": a, "PYRSGM_PATH_FORM": f, "PYRSGM_PATH_QUEUED": q})
    for m in (2, 1, 0) for a in (0, 1) for f in (0, 1, 2) for q in (0, 1) if not (m == 0 and f == 0)
]

SOURCES = PATH_PARTS + [
    "kernels_wta.hip",
    "kernels_cost.hip",
    "kernels_post.hip",
    "kernels_io.hip",
    "noalloc.hip",
    "capi.cpp",
    "stereo_matcher.cpp",
    "matcher_capi.cpp",
    "server.cpp",
    "mixed.cpp",
    "io.cpp",
    "prof.cpp",
]

ARCH = os.environ.get("PYRSGM_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def common_flags() -> list[str]:
    extra = os.environ.get("PYRSGM_EXTRA_FLAGS", "").split()
    return extra + [
        "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
        # Float expressions evaluated operation by operation (confidence, P2 table).
        "-ffp-contract=off",
        # No packed-FP32 instructions anywhere.
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-command-line-argument",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    ]


def _deps_mtime() -> float:
    hdrs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    inc = os.path.join(ROOT, "include", "pyrsgm")
    hdrs += [os.path.join(inc, f) for f in os.listdir(inc) if f.endswith(".h")]
    return max(os.path.getmtime(h) for h in hdrs)


def _compile(entry, force: bool = True) -> str:
    """entry: a source file name, or (source, object stem, {macro: value})."""
    if isinstance(entry, str):
        src, stem, defines = entry, os.path.splitext(entry)[0], {}
    else:
        src, stem, defines = entry
    obj = os.path.join(OBJDIR, stem + ".o")
    if not force and os.path.exists(obj):
        newest = max(os.path.getmtime(os.path.join(CSRC, src)), _deps_mtime())
        if os.path.getmtime(obj) > newest:
            return obj
    cmd = [hipcc(), *common_flags(), *[f"-D{k}={v}" for k, v in defines.items()],
           "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd.insert(1, "-x")
        cmd.insert(2, "hip")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        noise = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if noise:
            sys.stderr.write("\n".join(noise) + "\n")
    return obj


# On gfx950 a reversed-operand VOP2 instruction with a DPP lane swizzle
# (v_subrev_u32_dpp and the like) applies the swizzle to its second source,
# while the compiler's DPP combiner assumes the first (measured: v_subrev_u32_dpp
# with row_shr:1 returns y[l-1] - x[l], not y[l] - x[l-1]).  The combiner forms
# such instructions when it folds a DPP move whose out-of-row lanes read 0 into
# a subtraction or shift, so every build is checked for them.
_REV_DPP = re.compile(r"^\s+(v_[a-z0-9_]*rev[a-z0-9_]*_dpp)\s", re.M)


def _llvm_tool(name: str) -> str:
    for cand in (os.path.join(os.path.dirname(os.path.realpath(hipcc())), "..", "llvm", "bin", name),
                 f"/opt/rocm/llvm/bin/{name}", shutil.which(name)):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError(f"{name} not found")


def check_no_reversed_dpp(obj: str) -> None:
    """Raises if the gfx950 code of `obj` holds a reversed-operand DPP instruction."""
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        fat, dev = os.path.join(td, "fat.bin"), os.path.join(td, "dev.co")
        r = subprocess.run(["objcopy", f"--dump-section=.hip_fatbin={fat}", obj, os.path.join(td, "o")],
                           capture_output=True, text=True)
        if r.returncode != 0 or not os.path.exists(fat):
            return  # no device code in this object
        subprocess.run([_llvm_tool("clang-offload-bundler"), "--unbundle", "--type=o", f"--input={fat}",
                        f"--targets=hipv4-amdgcn-amd-amdhsa--{ARCH}", f"--output={dev}"],
                       check=True, capture_output=True)
        dis = subprocess.run([_llvm_tool("llvm-objdump"), "-d", f"--mcpu={ARCH}", dev],
                             check=True, capture_output=True, text=True).stdout
    bad = sorted(set(_REV_DPP.findall(dis)))
    if bad:
        raise RuntimeError(f"{os.path.basename(obj)}: reversed-operand DPP instructions {bad} "
                           "(their swizzle applies to the second source on gfx950); keep the DPP "
                           "move out of the combiner (see dppOpaque in device_common.h)")


def build(jobs: int | None = None, force: bool = True) -> str:
    """Compile every source (force) or only out-of-date ones, then link."""
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
        list(ex.map(check_no_reversed_dpp, objs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force="--incremental" not in sys.argv))
