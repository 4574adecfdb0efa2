"""AddressSanitizer / UndefinedBehaviorSanitizer runs on the CPU (no GPU code
is built with a sanitizer): the library's host code (mesh producers and the
drivers' setup-file reader, tests/asan/host_check.cpp) and the CPU oracle
(oracle/_asan/libnbo.so from `make -C oracle asan`, loaded by the oracle
tests through NBO_LIB with the ASan runtime preloaded into Python)."""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="no host compiler")


def asan_runtime():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], stdout=subprocess.PIPE, text=True).stdout.strip()
    return p if p and Path(p).exists() else None


def test_host_code_under_asan(tmp_path):
    exe = tmp_path / "host_check"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-ffp-contract=off", f"-I{ROOT / 'include'}", f"-I{ROOT / 'nekbench_amd' / 'apps'}",
           str(ROOT / "tests" / "asan" / "host_check.cpp"), str(ROOT / "nekbench_amd" / "csrc" / "nbx_mesh.cpp"),
           "-o", str(exe)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([str(exe), str(ROOT / "nekbench_amd" / "apps" / "setup.ini")], env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0 and "host_check: ok" in r.stdout, r.stdout[-4000:]


def test_oracle_under_asan():
    rt = asan_runtime()
    if rt is None:
        pytest.skip("no ASan runtime")
    r = subprocess.run(["make", "-C", str(ROOT / "oracle"), "-s", "asan"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    env = dict(os.environ, NBO_LIB=str(ROOT / "oracle" / "_asan" / "libnbo.so"), LD_PRELOAD=rt,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1",
               OMP_NUM_THREADS="4")
    tests = sorted(str(p) for p in (ROOT / "tests").glob("test_oracle_*.py"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *tests], env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=1200, cwd=str(ROOT))
    assert r.returncode == 0, r.stdout[-6000:]
    assert "AddressSanitizer" not in r.stdout and "runtime error" not in r.stdout
