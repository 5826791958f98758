"""CPU test of the serving loop's host-side policy: builds tests/cpp/policy_test.cpp
against sequoia_amd/csrc/server_policy.h with g++ (plain and with ASan/UBSan)
and runs it."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "policy_test.cpp")
INC = os.path.join(ROOT, "sequoia_amd", "csrc")


@pytest.mark.parametrize("flags", [["-O2"], ["-O1", "-g", "-fsanitize=address,undefined"]], ids=["O2", "asan"])
def test_policy(tmp_path, flags):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "policy_test")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", *flags, "-I", INC, SRC, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "policy tests passed" in r.stdout
