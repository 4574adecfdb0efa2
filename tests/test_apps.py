"""The command-line drivers bin/axhelm and bin/nekBone without a GPU: argument
handling (axhelm/main.cpp:28-59, nekBone/main.cpp:44-50) and the reading of
nekBone setup files with setupAide's rules (core/setupAide.cpp:83-243)
resolved into the settings of nekBone/main.cpp and setup.cpp:36-60."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "nekbench_amd" / "bin"
INI = ROOT / "nekbench_amd" / "apps" / "setup.ini"


def run(*args, timeout=60):
    return subprocess.run([str(a) for a in args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=timeout)


def settings(path):
    r = run(BIN / "nekBone", "--print-options", path)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, v = line.split("=", 1)
        out[k] = v
    return out


def write(tmp_path, text, name="case.ini"):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_binaries_built():
    for name in ("axhelm", "nekBone", "bw", "dot"):
        assert (BIN / name).is_file(), f"{name} not built (run __graft_entry__.build())"


def test_usage_messages():
    r = run(BIN / "axhelm", 7, 1, 10)
    assert r.returncode == 1 and r.stdout.startswith("Usage: ./axhelm N Ndim numElements")
    r = run(BIN / "nekBone")
    assert r.returncode == 1 and r.stdout.strip() == "usage: ./nekBone setupfile"
    r = run(BIN / "nekBone", ROOT / "no-such-file.ini")
    assert r.returncode == 1 and "cannot read" in r.stderr
    r = run(BIN / "bw")
    assert r.returncode == 1 and r.stdout.strip() == "Usage: ./bw SERIAL|CUDA|OPENCL"
    r = run(BIN / "dot", 7, 10, "HIP")
    assert r.returncode == 1 and r.stdout.startswith("Usage: ./dot N numElements")


def test_axhelm_rejects_bad_sizes():
    for args in ((0, 1, 10), (16, 1, 10), (7, 2, 10), (7, 1, 0)):
        r = run(BIN / "axhelm", *args, "HIP", "gfx950")
        assert r.returncode == 1 and "need 1 <= N <= 15" in r.stderr


def test_shipped_setup_file():
    """The readme configuration: BPMODE forces COPY and lambda1 = 0 (setup.cpp:43-53)."""
    s = settings(INI)
    assert (s["N"], s["NX"], s["NY"], s["NZ"], s["Nfields"]) == ("7", "16", "16", "32", "1")
    assert s["BPid"] == "5" and float(s["lambda1"]) == 0.0 and s["precon"] == "COPY"
    assert float(s["tol"]) == 1e-6 and s["Ntests"] == "1" and s["maxIterFixed"] == "5000"
    assert s["fixed"] == "1" and s["flexible"] == "0" and s["fused"] == "1"


def test_parser_rules(tmp_path):
    text = """leading text before any key is ignored
[POLYNOMIAL DEGREE]   5   # trailing comment
[BOX NX] 4 [BOX NY]
  5
[BOX NZ] /* block
comment */ 6
[BPMODE]
FALSE
[LAMBDA]
  2.5e-1
[PRECONDITIONER]
#COPY
JACOBI
[KRYLOV SOLVER]
PCG+FLEXIBLE
[THREAD MODEL]
NATIVE+SERIAL
[NREPETITIONS]
3
[FIXED ITERATION COUNT]
FALSE
[SOLVER TOLERANCE]
1e-9
[ARCH]
"quoted value [kept]"
"""
    s = settings(write(tmp_path, text))
    assert (s["N"], s["NX"], s["NY"], s["NZ"]) == ("5", "4", "5", "6")
    assert s["BPid"] == "0" and float(s["lambda1"]) == 0.25
    assert s["precon"] == "JACOBI" and s["flexible"] == "1"
    assert s["Ntests"] == "3" and s["fixed"] == "0" and float(s["tol"]) == 1e-9


def test_defaults_and_bp6(tmp_path):
    s = settings(write(tmp_path, "[POLYNOMIAL DEGREE]\n3\n"))
    # BOX N* default 3 (core/meshSetup.cpp), LAMBDA 1.1 (setup.cpp:51), tol 1e-8 and
    # NREPETITIONS 10 (main.cpp:96-103), 1000 iterations (main.cpp:14)
    assert (s["NX"], s["NY"], s["NZ"]) == ("3", "3", "3")
    assert float(s["lambda1"]) == 1.1 and float(s["tol"]) == 1e-8 and s["Ntests"] == "10"
    assert s["maxIterFixed"] == "1000" and s["precon"] == "COPY" and s["BPid"] == "0"
    s = settings(write(tmp_path, "[POLYNOMIAL DEGREE]\n7\n[NFIELDS]\n3\n[BPMODE]\nTRUE\n[PRECONDITIONER]\nJACOBI\n"))
    assert s["BPid"] == "6" and s["precon"] == "COPY" and s["Nfields"] == "3"


def test_multigrid_extension(tmp_path):
    s = settings(write(tmp_path, "[POLYNOMIAL DEGREE]\n7\n[PRECONDITIONER]\nMULTIGRID\n"))
    assert s["precon"] == "MULTIGRID" and s["mgOrders"] == "7,3,1" and s["mgDegree"] == "3"
    s = settings(write(tmp_path, "[POLYNOMIAL DEGREE]\n7\n[PRECONDITIONER]\nMULTIGRID\n[MULTIGRID ORDERS]\n7, 3, 1\n"
                                 "[MULTIGRID PRECISION]\nBF16\n[KERNEL SEQUENCE]\nREFERENCE\n"))
    assert s["mgOrders"] == "7,3,1" and s["mgPrec"] == "2" and s["fused"] == "0"


@pytest.mark.parametrize("text,msg", [
    ("[BOX NX]\n4\n", "missing [POLYNOMIAL DEGREE]"),
    ("[POLYNOMIAL DEGREE]\n16\n", "POLYNOMIAL DEGREE must be in [1, 15]"),
    ("[POLYNOMIAL DEGREE]\n7\n[BOX NX]\n2\n", "at least 3 elements"),
    ("[POLYNOMIAL DEGREE]\n7\n[NFIELDS]\n2\n", "NFIELDS must be 1 or 3"),
    ("[POLYNOMIAL DEGREE]\n7\n[KRYLOV SOLVER]\nGMRES\n", "only the PCG"),
])
def test_setup_errors(tmp_path, text, msg):
    r = run(BIN / "nekBone", "--print-options", write(tmp_path, text))
    assert r.returncode == 1 and msg in r.stderr
