"""The native library's ABI and host-side logic, on the CPU (no GPU calls).

* every C function declared in include/pyrsgm/*.h is exported by libpyrsgm.so,
  and so is the C++ API of pyrsgm.h;
* pyrsgm::validateParams follows the reference's own validation test
  (tests/matcher_test.cpp:41-72 of gegeges/pyrsgm) plus this build's block-size limit;
* the C entry points reject bad arguments before any device call (pointers
  passed here are never dereferenced: each call returns at its argument check).
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import sys

import pytest

import pyrsgm_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include", "pyrsgm")
FAKE = C.c_void_p(0x1000)  # non-null, never dereferenced


def declared_c_functions():
    """Names of the functions declared in the C headers (extern "C" blocks)."""
    names = set()
    for h in ("pyrsgm_c.h", "pyrsgm_py.h", "pyrsgm_prof.h", "pyrsgm_batch.h", "pyrsgm_serve.h",
              "pyrsgm_mixed.h", "pyrsgm_io.h"):
        text = open(os.path.join(INC, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(pyrsgm_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_reference_c_abi():
    names = declared_c_functions()
    for ref in ("pyrsgm_create_workspace", "pyrsgm_destroy_workspace", "pyrsgm_workspace_reserve",
                "pyrsgm_compute_single_scale", "pyrsgm_compute_level", "pyrsgm_pyrdown_u8"):
        assert ref in names


def test_library_exports_every_declared_c_function():
    lib = P.lib()
    missing = [n for n in declared_c_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_exports_the_cpp_api():
    out = subprocess.run(["nm", "-DC", "--defined-only", P.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    # invalidDisparity16 is an inline header function (reference pyrsgm.h:89-92).
    for sym in ("pyrsgm::validateParams(pyrsgm::Params const&)",
                "pyrsgm::StereoMatcher::StereoMatcher(int, int, pyrsgm::Params const&)",
                "pyrsgm::StereoMatcher::StereoMatcher(pyrsgm::StereoMatcher&&)",
                "pyrsgm::StereoMatcher::operator=(pyrsgm::StereoMatcher&&)",
                "pyrsgm::StereoMatcher::width() const",
                "pyrsgm::StereoMatcher::height() const",
                "pyrsgm::StereoMatcher::params() const",
                "pyrsgm::StereoMatcher::compute(",
                "pyrsgm::StereoMatcher::levelInfo(int) const",
                "pyrsgm::StereoMatcher::numLevels() const",
                "pyrsgm::StereoMatcher::deviceMemoryBytes() const",
                "pyrsgm::BatchMatcher::BatchMatcher(int, int, int, pyrsgm::Params const&)",
                "pyrsgm::BatchMatcher::compute(",
                "pyrsgm::BatchMatcher::capture(",
                "pyrsgm::BatchMatcher::launch(",
                "pyrsgm::BatchMatcher::batch() const",
                "pyrsgm::BatchMatcher::requiredBytes(int, int, int, pyrsgm::Params const&)",
                "pyrsgm::BatchMatcher::tune(",
                "pyrsgm::BatchMatcher::setTuning(pyrsgm::BatchTuning const&)",
                "pyrsgm::MixedBatch::MixedBatch(",
                "pyrsgm::MixedBatch::compute(pyrsgm_pair_io_c const*, ihipStream_t*)",
                "pyrsgm::MixedBatch::capture(pyrsgm_pair_io_c const*)",
                "pyrsgm::MixedBatch::launch(ihipStream_t*)",
                "pyrsgm::MixedBatch::tune(",
                "pyrsgm::Server::Server(pyrsgm::ServeOptions const&)",
                "pyrsgm::Server::submit(pyrsgm::ServeRequest const&, long)",
                "pyrsgm::Server::wait(unsigned long)",
                "pyrsgm::Server::poll(unsigned long, pyrsgm_serve_result_c*)",
                "pyrsgm::Server::drain()",
                "pyrsgm::Server::stats() const",
                "pyrsgm::version()"):
        assert sym in out, sym


def test_batch_rejects_bad_sizes_on_cpu():
    # Validation happens before any device allocation.
    with pytest.raises(ValueError):
        P.BatchMatcher(64, 32, 0, P.Params(num_disparities=64))
    with pytest.raises(ValueError):
        P.BatchMatcher(0, 32, 2, P.Params(num_disparities=64))
    with pytest.raises(ValueError):
        P.BatchMatcher(64, 32, 2, P.Params(p1=10, p2=5))


def test_mixed_rejects_bad_pairs_on_cpu():
    # Pair list and parameter checks run before any device call.
    with pytest.raises(ValueError):
        P.MixedBatch([])
    with pytest.raises(ValueError):
        P.MixedBatch([(0, 32, P.Params(num_disparities=64))])
    with pytest.raises(ValueError):
        P.MixedBatch([(64, 32, P.Params(num_disparities=64)), (64, 32, P.Params(p1=10, p2=5))])
    L = P.lib()
    assert L.pyrsgm_mixed_compute(None, None, None, None, 0) == -1
    assert L.pyrsgm_mixed_launch(None, None, None, 0) == -1
    assert L.pyrsgm_mixed_device_bytes(None) == 0


def test_batch_required_bytes_on_cpu():
    # The allocation formula needs no device: it grows with the pair count and
    # validates its arguments.
    p = P.Params(num_disparities=128, pyramid_levels=2, cost_mode=P.CostMode.Census)
    one = P.BatchMatcher.required_bytes(960, 540, 1, p)
    assert one > 960 * 540 * 16 * 2
    assert P.BatchMatcher.required_bytes(960, 540, 4, p) >= 4 * one - 4 * 256 * 540
    with pytest.raises(ValueError):
        P.BatchMatcher.required_bytes(960, 540, 0, p)
    with pytest.raises(ValueError):
        P.BatchMatcher.required_bytes(960, 540, 1, P.Params(p1=10, p2=5))


def test_server_rejects_bad_options_on_cpu():
    # Option checks run before any device call.
    with pytest.raises(ValueError):
        P.Server(max_batch=0)
    with pytest.raises(ValueError):
        P.Server(lanes_per_key=0)


def test_version():
    assert P.version() == "0.1.0"


def throws(**kw):
    with pytest.raises(ValueError):
        P.validate_params(P.Params(**kw))


def test_validation_follows_reference_test():
    v = P.validate_params(P.Params())
    assert v.num_disparities == 128 and v.block_size == 5
    assert v.pyramid_max_disparity == 0
    assert P.validate_params(P.Params(num_disparities=100)).num_disparities == 112
    assert P.validate_params(P.Params(num_disparities=100, block_size=4)).block_size == 5
    assert P.validate_params(P.Params(pyramid_levels=2)).pyramid_max_disparity == 127
    throws(pyramid_levels=5)
    throws(pyramid_levels=1, min_disparity=8)
    throws(pyramid_levels=1, disp12_max_diff=-1)
    throws(num_disparities=2048)
    throws(p1=100, p2=50)
    throws(adaptive_p2=True, p2_beta=0.0)


def test_validation_other_reference_rules():
    # stereo_matcher.cpp:100-148
    throws(num_disparities=0)
    assert P.validate_params(P.Params(num_disparities=1020)).num_disparities == 1024
    throws(block_size=0)
    assert P.validate_params(P.Params(block_size=1)).block_size == 3
    throws(p1=-1, p2=10)
    throws(uniqueness_ratio=100)
    throws(uniqueness_ratio=-1)
    throws(pre_filter_cap=0)
    assert P.validate_params(P.Params(num_paths=7)).num_paths == 3
    assert P.validate_params(P.Params(num_paths=0)).num_paths == 1
    throws(pyramid_levels=-1)
    throws(pyramid_levels=1, pyramid_coarse_disparities=0)
    v = P.validate_params(P.Params(pyramid_levels=1, pyramid_coarse_disparities=1500))
    assert v.pyramid_coarse_disparities == 1024
    assert P.validate_params(P.Params(pyramid_levels=1, pyramid_coarse_disparities=20)).pyramid_coarse_disparities == 32
    v = P.validate_params(P.Params(pyramid_levels=1, pyramid_min_confidence=3.0))
    assert v.pyramid_min_confidence == 1.0
    # Legacy mode leaves the pyramid fields alone.
    assert P.validate_params(P.Params(pyramid_min_confidence=3.0)).pyramid_min_confidence == 3.0


def test_validation_block_size_limit():
    # This build's limit (the BT window sum and the right-view key assume it).
    assert P.validate_params(P.Params(block_size=255)).block_size == 255
    throws(block_size=256)


def test_invalid_disparity16():
    assert P.invalid_disparity16(P.Params()) == -16
    assert P.invalid_disparity16(P.Params(min_disparity=5)) == 64
    assert P.invalid_disparity16(P.Params(min_disparity=-3)) == -64


def single(ws, w=64, h=32, D=16, block=5, P1=8, P2=32, left=FAKE, right=FAKE):
    return P.compute_single_scale(ws, left, right, FAKE, FAKE, FAKE, w, w, 2 * w, 4 * w, 4 * w,
                                  w, h, D, 0, block, P1, P2, 10, 1, 63, 1, 3, 3, 0,
                                  3.041, 16.137, 0.601, 1.77, None)


def level(ws, w=64, h=32, D=16, maxD=15, refine=0, parent=FAKE, pw=32, ph=16, block=5, P1=8, P2=32):
    return P.compute_level(ws, FAKE, FAKE, parent, parent, FAKE, FAKE, FAKE, w, w, 2 * pw, 4 * pw,
                           2 * w, 4 * w, 4 * w, w, h, pw, ph, D, 0, maxD, block, P1, P2, 10, 1,
                           63, 1, 1, 3, refine, 0.05, 0, 3.041, 16.137, 0.601, 1.77, None)


@pytest.fixture()
def ws():
    w = P.create_workspace()
    assert w
    yield w
    P.destroy_workspace(w)  # no device memory was allocated


def test_c_abi_single_scale_rejections(ws):
    assert single(None) == -1
    assert single(ws, left=None) == -1
    assert single(ws, right=None) == -1
    assert single(ws, w=0) == -1
    assert single(ws, h=0) == -1
    assert single(ws, w=70000) == -1          # right-view key packs the column in 16 bits
    assert single(ws, D=20) == -1             # reference: D % 16 == 0 (sgm_kernels.cu:1610)
    assert single(ws, D=0) == -1
    assert single(ws, D=1040) == -1           # D <= 1024
    assert single(ws, block=0) == -1
    assert single(ws, block=256) == -1
    assert single(ws, P1=-1) == -1
    assert single(ws, P2=-1) == -1


def test_c_abi_level_rejections(ws):
    assert level(None) == -1
    assert level(ws, D=24) == -1
    assert level(ws, D=272, maxD=271, refine=1) == -1   # refinement D <= 256 (sgm_kernels.cu:1651)
    assert level(ws, refine=1, parent=None) == -1        # (sgm_kernels.cu:1654-1657)
    assert level(ws, refine=1, pw=0) == -1
    assert level(ws, refine=1, ph=0) == -1
    assert level(ws, maxD=-1) == -1                       # maxD < minD after clamping
    assert level(ws, P2=-5) == -1


def test_c_abi_reserve_and_pyrdown_rejections(ws):
    assert P.workspace_reserve(ws, 0, 10, 16) == -1
    assert P.workspace_reserve(ws, 10, 10, 17) == -1
    assert P.workspace_reserve(None, 10, 10, 16) == -1
    # pyrdown: dst must be ((w+1)/2, (h+1)/2), src at least 2x2 (pyrdown.cu:56-74)
    assert P.pyrdown_u8(FAKE, 10, 10, 10, FAKE, 5, 4, 5, None) == -1
    assert P.pyrdown_u8(FAKE, 10, 10, 10, FAKE, 5, 5, 6, None) == -1
    assert P.pyrdown_u8(FAKE, 1, 1, 10, FAKE, 1, 1, 5, None) == -1
    assert P.pyrdown_u8(None, 10, 10, 10, FAKE, 5, 5, 5, None) == -1
    assert P.pyrdown_u8(FAKE, 10, 10, 10, None, 5, 5, 5, None) == -1


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
