"""CPU runs under AddressSanitizer and UndefinedBehaviorSanitizer:

* the scene-file code (sequoia_amd/csrc/sceneio.cpp, built alone with g++)
  driven by tests/cpp/sceneio_fuzz.cpp: seeded mutations of the OBJ grammar
  file tests/golden/obj/quirks.obj through sq_load_obj, and random image sizes
  and short buffers through the BMP encoder;
* every entry point of the oracle (oracle/*.cpp) on small seeded scenes,
  driven by tests/cpp/oracle_asan.cpp.

Any out-of-bounds access, leak or undefined behaviour makes the driver exit
non-zero. (The serving loop's arena and admission policy run under the same
sanitizers in test_server_policy.py.)"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
           OMP_NUM_THREADS="4")


def _need_gxx():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, env=ENV, **kw)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    return r.stdout


def test_sceneio_fuzz(tmp_path):
    _need_gxx()
    exe = str(tmp_path / "sceneio_fuzz")
    _run(["g++", "-std=c++17", "-Wall", "-Werror", *SAN, "-I", os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "cpp", "sceneio_fuzz.cpp"),
          os.path.join(ROOT, "sequoia_amd", "csrc", "sceneio.cpp"), "-o", exe])
    out = _run([exe, os.path.join(ROOT, "tests", "golden", "obj", "quirks.obj"), str(tmp_path), "20000"],
               timeout=300)
    # both outcomes are exercised: files that load and files that are refused
    words = out.split()
    loaded, rejected = int(words[words.index("loaded,") - 1]), int(words[words.index("rejected") - 1])
    assert loaded > 1000 and rejected > 1000, out


def test_oracle_under_sanitizers(tmp_path):
    _need_gxx()
    exe = str(tmp_path / "oracle_asan")
    od = os.path.join(ROOT, "oracle")
    _run(["g++", "-std=c++17", "-Wall", "-Werror", "-Wno-unused-parameter", "-fopenmp", "-ffp-contract=off", *SAN,
          "-I", od, os.path.join(ROOT, "tests", "cpp", "oracle_asan.cpp"),
          *(os.path.join(od, f) for f in ("oracle.cpp", "node2.cpp", "nodek.cpp", "plucker.cpp")), "-o", exe])
    out = _run([exe], timeout=300)
    assert "trees built and traced" in out, out
