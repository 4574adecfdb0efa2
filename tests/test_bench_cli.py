"""bench.py's host-side contract (no GPU): a run on more than one GPU is
refused with a message, a solve that stopped at its cap gets no speedup, the
flat time-to-tolerance summary picks the fastest converged solver, and the
cpu_baseline record states its cores, kind and sample."""
import importlib.util
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120, cwd=str(ROOT))
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_capped_solves_get_no_speedup():
    b = load_bench()
    out = {"fp64_jacobi": {"time_s": 0.6, "converged": True},
           "fast_but_capped": {"time_s": 0.01, "converged": False},
           "mg": {"time_s": 0.03, "converged": True}}
    b.set_speedups(out)
    assert out["fp64_jacobi"]["speedup_vs_fp64_jacobi"] == 1.0
    assert out["fast_but_capped"]["speedup_vs_fp64_jacobi"] is None
    assert out["mg"]["speedup_vs_fp64_jacobi"] == 20.0
    flat = b.flat_time_to_tol(out, "t")
    assert flat["t_best_solver"] == "mg" and flat["t_best_s"] == 0.03 and flat["t_best_speedup"] == 20.0
    assert flat["t_fp64_jacobi_s"] == 0.6
    # the reference solve itself capped: no speedup for anyone, no Jacobi time
    out2 = {"fp64_jacobi": {"time_s": 0.6, "converged": False}, "mg": {"time_s": 0.03, "converged": True}}
    b.set_speedups(out2)
    assert out2["mg"]["speedup_vs_fp64_jacobi"] is None
    assert b.flat_time_to_tol(out2, "t")["t_fp64_jacobi_s"] is None


def test_learned_entries_are_flattened():
    b = load_bench()
    out = {"fp64_jacobi": {"time_s": 0.6, "converged": True},
           "learned_mixed_mg_bf16cycle": {"time_s": 0.032, "converged": True, "gain_vs_same_mg": 0.9}}
    b.set_speedups(out)
    flat = b.flat_time_to_tol(out, "x")
    assert flat["x_learned_mixed_mg_bf16cycle_s"] == 0.032
    assert flat["x_learned_mixed_mg_bf16cycle_gain_vs_same_mg"] == 0.9
    assert abs(flat["x_learned_mixed_mg_bf16cycle_speedup"] - 18.75) < 1e-9


def test_cpu_baseline_record():
    b = load_bench()
    res = {"cores": {"threads": 16, "nproc": 16, "physical_cores_in_affinity": 16, "cpu": "AMD EPYC"},
           "build": "gcc -O3", "pcg": {"value": 0.6, "s_per_iteration": 0.0187, "k1": 4, "k2": 24,
                                        "elements": 32768},
           "axhelm": {"kind": "reference", "us_per_apply": 554.0, "oracle_port_us_per_apply": 489.0,
                      "oracle_port_over_reference": 0.883}}
    rec = b.cpu_baseline_record(res)
    assert rec["value"] == 0.6 and rec["unit"] == "GDOF/s" and rec["cores"] == 16 and rec["kind"] == "port"
    assert "32^3" in rec["sample"] and "16 physical cores" in rec["sample"]
    assert rec["calibration"]["port_over_reference"] == 0.883
    assert b.cpu_baseline_record(None) is None
