"""Bandwidth probes of the bw and dot benchmarks (bw/kernel/bwKernels.okl,
dot/main.cpp) and their drivers.  The triad and the LDS probe are compared
bit for bit with numpy evaluations in the same order; the read stream's sum
is deterministic and within rounding of an exactly rounded sum."""
import math
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import nekbench_amd as nbx

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "nekbench_amd" / "bin"

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.mark.parametrize("n", [1, 2, 3, 1001, 1 << 20, (1 << 20) + 7])
def test_triad_copy_read(torch, n):
    print(f"[bw] triad/copy/read n={n}", flush=True)
    L = nbx.lib()
    rng = np.random.default_rng(n)
    a_h, b_h = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    a, b = torch.from_numpy(a_h).cuda(), torch.from_numpy(b_h).cuda()
    c = torch.empty(n, dtype=torch.float64, device="cuda")
    assert L.nbx_triad(n, 0.75, a.data_ptr(), b.data_ptr(), c.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(c.cpu().numpy(), a_h + 0.75 * b_h)
    d = torch.empty_like(a)
    assert L.nbx_copy(n, a.data_ptr(), d.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), a_h)
    w1 = torch.full((2049,), float("nan"), dtype=torch.float64, device="cuda")
    w2 = torch.zeros(2049, dtype=torch.float64, device="cuda")
    assert L.nbx_read_sum(n, a.data_ptr(), w1.data_ptr(), None) == 0
    assert L.nbx_read_sum(n, a.data_ptr(), w2.data_ptr(), None) == 0
    torch.cuda.synchronize()
    s1, s2 = float(w1[0]), float(w2[0])
    assert s1 == s2
    exact = math.fsum(a_h)
    assert abs(s1 - exact) <= 1e-13 * np.abs(a_h).sum()


def test_alignment_and_argument_checks(torch):
    L = nbx.lib()
    a = torch.zeros(16, dtype=torch.float64, device="cuda")
    off = a.data_ptr() + 8
    w = torch.zeros(2049, dtype=torch.float64, device="cuda")
    assert L.nbx_triad(4, 1.0, off, a.data_ptr(), a.data_ptr(), None) == -1
    assert L.nbx_copy(4, a.data_ptr(), off, None) == -1
    assert L.nbx_read_sum(4, off, w.data_ptr(), None) == -1
    assert L.nbx_read_sum(0, a.data_ptr(), w.data_ptr(), None) == -1
    assert L.nbx_read_sum(4, a.data_ptr(), None, None) == -1
    assert L.nbx_lds_probe(-1, a.data_ptr(), None) == -1
    assert L.nbx_triad(0, 1.0, a.data_ptr(), a.data_ptr(), a.data_ptr(), None) == 0


def test_lds_probe(torch):
    L = nbx.lib()
    nb = 300
    a_h = np.random.default_rng(7).uniform(-1, 1, nb * 128)
    a = torch.from_numpy(a_h).cuda()
    assert L.nbx_lds_probe(nb, a.data_ptr(), None) == 0
    torch.cuda.synchronize()
    blocks = a_h.reshape(nb, 128).copy()
    x = blocks[:, :1].copy()
    for _ in range(128):
        blocks = blocks + x
    assert np.array_equal(a.cpu().numpy(), blocks.ravel())


def test_bw_driver():
    print("[driver] bw HIP", flush=True)
    r = subprocess.run([str(BIN / "bw"), "HIP"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert len(re.findall(r"^triad stream: \d+ words", out, re.M)) == 12
    assert "shared mem bw:" in out and "D->H memcpy" in out and "H->D memcpy" in out and "D->D memcpy" in out
    for name in ("triad", "copy", "read"):
        m = re.search(rf"^HBM {name} stream: \d+ words, \S+ s, (\S+) bytes/s", out, re.M)
        assert m and 1e12 < float(m.group(1)) < 8.5e12, out


def test_dot_driver():
    print("[driver] dot 7 4096 HIP gfx950", flush=True)
    r = subprocess.run([str(BIN / "dot"), "7", "4096", "HIP", "gfx950", "20"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    m = re.search(r"Correctness check: \|wxy - host sum\| / \|host sum\| = (\S+)", r.stdout)
    assert m and float(m.group(1)) < 1e-13
    assert re.search(r"^MPItasks=1 .* N=7 Nelements=4096 blockSize=256 elapsed time=\S+ GDOF/s=\S+ GB/s=\S+$",
                     r.stdout, re.M), r.stdout
