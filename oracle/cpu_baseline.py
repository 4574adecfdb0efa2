"""CPU baseline of the benchmark, run as a child process of bench.py so that its
OpenMP placement (OMP_PLACES=cores, OMP_PROC_BIND=close) binds only this
process's threads and never the threads that drive the GPU.

TEST INFRASTRUCTURE ONLY: the CPU oracle (libnbo.so) and the reference's own
CPU kernel compiled into oracle/_ref are the checker and the baseline, never
the product.  Nothing here touches a GPU.

  --pcg NX      FP64 Jacobi PCG of the NX^3-element N=7 perturbed Dirichlet
                box (the benchmark's problem, oracle restatement of
                nekBone/solve.cpp:44-113): per-iteration time as the
                difference of two fixed-count solves (k2 - k1 iterations), so
                the initial residual r = b - A x and the Jacobi setup that
                both solves do cancel; one warm-up solve first, then the
                fastest of three timings of each count.
  --axhelm NX   FP64 BK axhelm on the NX^3-element N=7 perturbed box:
                the reference's axhelm_bk_v0 (axhelm/kernel/CPU/axhelm.c) and
                the oracle's restatement, each one warm-up call and then the
                mean of 100 calls (axhelm/main.cpp:132-142).
Prints one JSON object on stdout.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def core_info():
    """Threads the oracle runs (OpenMP), logical CPUs of the machine and of
    this process's affinity set, and the physical cores behind that set."""
    import nbo
    aff = sorted(os.sched_getaffinity(0))
    phys, cur = {}, {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if "processor" in cur:
                    phys[cur["processor"]] = (cur.get("physical id", "0"), cur.get("core id", cur["processor"]))
                cur = {}
                continue
            k, v = (s.strip() for s in line.split(":", 1))
            cur[k] = v
        if "processor" in cur:
            phys[cur["processor"]] = (cur.get("physical id", "0"), cur.get("core id", cur["processor"]))
    except OSError:
        pass
    cores = {phys.get(str(c), ("?", str(c))) for c in aff}
    return {"threads": int(nbo.lib().nbo_max_threads()), "nproc": os.cpu_count(), "affinity_cpus": len(aff),
            "physical_cores_in_affinity": len(cores), "cpu": cpu_model(),
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PLACES", "OMP_PROC_BIND")}}


def pcg_per_iteration(NX, k1, k2):
    import numpy as np
    import nbo
    N = 7
    box = nbo.Box.make(N, NX, NX, NX, perturb=0.1, seed=1)
    t0 = time.perf_counter()
    ob = nbo.OracleBP(box, bpmode=True)
    t_setup = time.perf_counter() - t0

    def solve(k):
        r, x = ob.rhs.copy(), np.zeros(ob.offset)
        t = time.perf_counter()
        ob.pcg(r, x, 0.0, k, jacobi=True)
        return time.perf_counter() - t

    solve(1)  # warm-up
    # the fixed part of a solve (buffers first touched, Jacobi diagonal,
    # r = b - A x) varies from run to run: each count is timed three times,
    # interleaved, and the fastest of each is kept
    t1, t2 = float("inf"), float("inf")
    for _ in range(3):
        t1 = min(t1, solve(k1))
        t2 = min(t2, solve(k2))
    t_it = (t2 - t1) / (k2 - k1)
    return {"value": box.E * N ** 3 / t_it / 1e9, "unit": "GDOF/s", "s_per_iteration": t_it,
            "k1": k1, "k2": k2, "t_k1_s": t1, "t_k2_s": t2, "setup_s": t_setup, "elements": box.E}


def axhelm_calibration(NX, reps):
    import numpy as np
    import nbo
    box = nbo.Box.make(7, NX, NX, NX, perturb=0.1, seed=1)
    m = nbo.box_mesh(box)
    E, n = box.E, box.E * box.Np
    q = np.random.default_rng(1).uniform(-1, 1, n)
    out = np.zeros(n)

    def mean_time(fn):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t0) / reps

    res = {"elements": E, "reps": reps}
    t_port = mean_time(lambda: nbo.axhelm(7, 1, 1, E, n, m["ggeo"], m["D"], None, q, out=out))
    res["oracle_port_us_per_apply"] = round(t_port * 1e6, 1)
    res["oracle_port_gdofs"] = round(E * 343 / t_port / 1e9, 3)
    if nbo.ref_available(7):
        t = mean_time(lambda: nbo.ref_axhelm(7, "axhelm_bk_v0", E, n, m["ggeo"], m["D"], None, q, out=out))
        res.update({"kind": "reference", "kernel": "axhelm_bk_v0 of axhelm/kernel/CPU/axhelm.c",
                    "us_per_apply": round(t * 1e6, 1), "GB_s": round(E * 512 * 72 / t / 1e9, 1),
                    "gdofs": round(E * 343 / t / 1e9, 3), "oracle_port_over_reference": round(t_port / t, 3)})
    else:
        res["kind"] = "port"
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pcg", type=int, default=0, help="NX of the PCG problem (0: skip)")
    ap.add_argument("--k1", type=int, default=4)
    ap.add_argument("--k2", type=int, default=24)
    ap.add_argument("--axhelm", type=int, default=0, help="NX of the axhelm mesh (0: skip)")
    ap.add_argument("--reps", type=int, default=100)
    a = ap.parse_args()
    out = {"cores": core_info(), "build": "gcc -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off (oracle/Makefile)"}
    if a.pcg:
        out["pcg"] = pcg_per_iteration(a.pcg, a.k1, a.k2)
    if a.axhelm:
        out["axhelm"] = axhelm_calibration(a.axhelm, a.reps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
