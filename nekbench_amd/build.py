"""Ahead-of-time build of libnbx.so (HIP kernels for gfx950 + host C++).

Every .hip file is compiled by hipcc for gfx950 only; host-only .cpp files by
g++.  FP contraction is off everywhere (the FP64 kernels spell out their
fused multiply-adds so that they reproduce the CPU oracle bit for bit) and
packed-FP32 instructions are disabled for device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
OBJ = PKG / "_build"
LIB = PKG / "libnbx.so"
APPS = PKG / "apps"
BIN = PKG / "bin"

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or f"{ROCM}/bin/hipcc"
ARCH = os.environ.get("NBX_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-mcode-object-version=5", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", f"-I{INCLUDE}",
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-Wall", f"-I{INCLUDE}"]


def _jobs() -> int:
    try:
        return max(1, min(16, int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    except ValueError:
        return 4


def _compile_cmds():
    cmds = []
    for src in sorted(CSRC.glob("*.hip")):
        obj = OBJ / (src.stem + ".hip.o")
        cmds.append((src, obj, [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]))
    for src in sorted(CSRC.glob("*.cpp")):
        obj = OBJ / (src.stem + ".cpp.o")
        if src.name == "nbx_core.cpp":
            cmd = [HIPCC, "-O3", "-fPIC", "-std=c++17", f"-I{INCLUDE}", "-c", str(src), "-o", str(obj)]
        else:
            cmd = ["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)]
        cmds.append((src, obj, cmd))
    return cmds


def build(verbose: bool = True) -> Path:
    """Compile every source and link libnbx.so; raises on any failure."""
    OBJ.mkdir(exist_ok=True)
    cmds = _compile_cmds()
    running = []
    pending = list(cmds)
    failures = []
    while pending or running:
        while pending and len(running) < _jobs():
            src, obj, cmd = pending.pop(0)
            if verbose:
                print(f"[nbx build] {src.name}", flush=True)
            running.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        src, proc = running.pop(0)
        out, _ = proc.communicate()
        if proc.returncode != 0:
            failures.append((src, out.decode(errors="replace")))
        elif verbose and out.strip():
            text = out.decode(errors="replace")
            lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and "target-feature" not in l]
            if any(l.strip() for l in lines):
                print("\n".join(lines), flush=True)
    if failures:
        msg = "\n".join(f"--- {s.name}\n{o}" for s, o in failures)
        raise RuntimeError(f"nbx build failed:\n{msg}")
    objs = [str(o) for _, o, _ in cmds]
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(LIB) + ".tmp", *objs, "-lgomp"]
    r = subprocess.run(link, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("nbx link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(str(LIB) + ".tmp", LIB)
    if verbose:
        print(f"[nbx build] linked {LIB}", flush=True)
    build_apps(verbose)
    build_native(verbose)
    return LIB


# command-line drivers with the reference's interfaces: bin/axhelm
# (axhelm/main.cpp), bin/nekBone (nekBone/main.cpp), bin/bw (bw/main.cpp) and
# bin/dot (dot/main.cpp); host C++ over libnbx.so
APP_SOURCES = {"axhelm": "axhelm_main.cpp", "nekBone": "nekbone_main.cpp", "bw": "bw_main.cpp",
               "dot": "dot_main.cpp"}


def build_apps(verbose: bool = True) -> None:
    BIN.mkdir(exist_ok=True)
    for name, src in APP_SOURCES.items():
        out = BIN / name
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{INCLUDE}",
               f"-I{APPS}", str(APPS / src), "-o", str(out) + ".tmp", f"-L{PKG}", "-lnbx", f"-L{ROCM}/lib",
               "-lamdhip64", "-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath,{ROCM}/lib"]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if r.returncode != 0:
            raise RuntimeError(f"{name} build failed:\n" + r.stdout.decode(errors="replace"))
        os.replace(str(out) + ".tmp", out)
        if verbose:
            print(f"[nbx build] {out}", flush=True)


NATIVE = PKG / "native"
NATIVE_ORDERS = range(1, 16)


def build_native(verbose: bool = True) -> None:
    """native/libaxhelm_N<N>.so for N = 1..15: the reference's native axhelm
    symbols (axhelm_bk_v0, ...) over libnbx.so, one library per order as the
    reference compiles its CPU kernel per p_Nq."""
    NATIVE.mkdir(exist_ok=True)
    src = APPS / "nbx_native_axhelm.cpp"
    procs = []
    for n in NATIVE_ORDERS:
        out = NATIVE / f"libaxhelm_N{n}.so"
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared", f"-DNBX_NATIVE_N={n}", f"-I{INCLUDE}",
               str(src), "-o", str(out) + ".tmp", f"-L{PKG}", "-lnbx", "-Wl,-rpath,$ORIGIN/..",
               f"-Wl,-rpath,{ROCM}/lib"]
        procs.append((out, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    for out, p in procs:
        text, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"{out.name} build failed:\n" + text.decode(errors="replace"))
        os.replace(str(out) + ".tmp", out)
    if verbose:
        print(f"[nbx build] {NATIVE}/libaxhelm_N1..15.so", flush=True)


def build_oracle(verbose: bool = True) -> None:
    """Build the CPU oracle and, when the nekBench tree is mounted, the
    reference CPU kernels under oracle/_ref (test infrastructure)."""
    odir = PKG.parent / "oracle"
    r = subprocess.run(["make", "-C", str(odir), f"-j{_jobs()}"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout.decode(errors="replace"))
    ref = Path(os.environ.get("NBX_REFERENCE", "/root/reference"))
    if (ref / "axhelm" / "kernel" / "CPU" / "axhelm.c").exists():
        r = subprocess.run(["make", "-C", str(odir), f"-j{_jobs()}", "ref", f"REF={ref}"],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if r.returncode != 0:
            raise RuntimeError("reference CPU kernel build failed:\n" + r.stdout.decode(errors="replace"))
    if verbose:
        print("[nbx build] oracle built", flush=True)


if __name__ == "__main__":
    build()
    build_oracle()
    sys.exit(0)
