"""The CMake package: the project configures, and (with PYRSGM_CMAKE_FULL=1, a
few minutes of compiling) builds, installs and is consumed through
find_package(pyrsgm) / pyrsgm::pyrsgm as the reference's README.md:154-155
describes."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CMAKE = shutil.which("cmake")
GEN = ["-G", "Ninja"] if shutil.which("ninja") else []

pytestmark = pytest.mark.skipif(
    CMAKE is None or not os.path.exists("/opt/rocm/llvm/bin/clang++"),
    reason="cmake or the ROCm compiler is not installed")


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}"
    return r


def test_configure(tmp_path):
    _run([CMAKE, "-S", ROOT, "-B", str(tmp_path / "b"), *GEN, "-DCMAKE_BUILD_TYPE=Release"])
    assert (tmp_path / "b" / "pyrsgmConfig.cmake").exists()
    assert (tmp_path / "b" / "pyrsgmConfigVersion.cmake").exists()


@pytest.mark.skipif(os.environ.get("PYRSGM_CMAKE_FULL") != "1",
                    reason="full CMake build takes minutes; set PYRSGM_CMAKE_FULL=1")
def test_build_install_consume(tmp_path):
    b, inst, cons = tmp_path / "b", tmp_path / "inst", tmp_path / "cons"
    _run([CMAKE, "-S", ROOT, "-B", str(b), *GEN, "-DCMAKE_BUILD_TYPE=Release"])
    _run([CMAKE, "--build", str(b), "-j", str(min(8, os.cpu_count() or 1))])
    _run([CMAKE, "--install", str(b), "--prefix", str(inst)])
    assert (inst / "lib" / "cmake" / "pyrsgm" / "pyrsgmTargets.cmake").exists()
    _run([CMAKE, "-S", os.path.join(ROOT, "tests", "cmake_consumer"), "-B", str(cons), *GEN,
          f"-DCMAKE_PREFIX_PATH={inst}"])
    _run([CMAKE, "--build", str(cons)])
    r = _run([str(cons / "consumer")])
    assert r.stdout.startswith("pyrsgm ")
