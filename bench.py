#!/usr/bin/env python3
"""Benchmark of the nekBench solver hot path on one MI355X.

Workload (one step = one PCG solve, inputs resident in HBM):
  32x32x32 = 32768 hex elements, N=7 (~11.1M unknowns), unit box with
  interior vertices perturbed by +-0.1 element width (seed 1), homogeneous
  Dirichlet boundary, rhs U[-1,1) per local node masked and assembled with
  ogs, FP64 Jacobi-preconditioned CG stopping at a 1e-8 relative weighted
  residual or 500 iterations (nekBone BPPCG, Poisson operator).

value = nekBone's throughput convention Nfields * iterations * E * N^3 / t
(GDOF/s, nekBone/main.cpp:156-158) with t the wall time of one solve.
roofline: the FP64 axhelm kernel of the solve (BK, N=7, 32768 elements),
algorithmic bytes E*Np*(2+7)*8 per launch (axhelm/main.cpp:162-163) over the
average launch time measured with HIP events on the library's stream;
traffic = HBM bytes per launch from the committed rocprofv3 FETCH_SIZE /
WRITE_SIZE runs (profiles/traffic.json, calibrated on a 1 GiB copy).
cpu_baseline: the CPU oracle's restatement of the same solve, timed in a
child process (oracle/cpu_baseline.py, OpenMP with OMP_PLACES=cores and
OMP_PROC_BIND=close on the host's CPU share): the per-iteration time as the
difference of two fixed-count solves, so the setup of each solve cancels.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "PCG time-to-1e-8 residual, learned/mixed vs FP64 Jacobi; axhelm GDOF/s"
HBM_PEAK_GBS = 8000.0
ROOF_SETS = 4  # rotating buffer sets of the roofline measurement


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_cpu_baseline(pcg_nx, axhelm_nx, log):
    """The CPU baseline in a child process (no GPU there): OpenMP placement
    binds that process's threads only.  Threads = OMP_NUM_THREADS when set,
    else the CPUs of this process's affinity set."""
    import subprocess
    env = dict(os.environ)
    env.setdefault("OMP_NUM_THREADS", str(len(os.sched_getaffinity(0))))
    env["OMP_PLACES"] = "cores"
    env["OMP_PROC_BIND"] = "close"
    cmd = [sys.executable, str(ROOT / "oracle" / "cpu_baseline.py")]
    if pcg_nx:
        cmd += ["--pcg", str(pcg_nx)]
    if axhelm_nx:
        cmd += ["--axhelm", str(axhelm_nx)]
    r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=900)
    if r.returncode != 0:
        log(f"[bench] cpu baseline failed ({r.returncode}): {r.stderr.decode(errors='replace')[-2000:]}")
        return None
    res = json.loads(r.stdout.decode().strip().splitlines()[-1])
    log(f"[bench] cpu baseline: {res}")
    return res


def cpu_baseline_record(res):
    """The benchmark line's cpu_baseline object from the child's result."""
    if not res or "pcg" not in res:
        return None
    p, c = res["pcg"], res["cores"]
    rec = {"value": p["value"], "unit": "GDOF/s", "cores": c["threads"], "kind": "port",
           "sample": f"CPU oracle restatement of the same {round(p['elements'] ** (1 / 3))}^3 N=7 Jacobi PCG "
                     f"(nbo_bp_pcg, OpenMP, {c['threads']} threads on {c['physical_cores_in_affinity']} physical "
                     f"cores of {c['cpu']}, nproc {c['nproc']}, OMP_PLACES=cores OMP_PROC_BIND=close): "
                     f"(t[{p['k2']} it] - t[{p['k1']} it]) / {p['k2'] - p['k1']} = {p['s_per_iteration']:.4f} s "
                     f"per iteration, setup cancelled; {res['build']}"}
    ax = res.get("axhelm")
    if ax and ax.get("kind") == "reference":
        rec["calibration"] = {"axhelm_reference_us": ax["us_per_apply"], "axhelm_port_us": ax["oracle_port_us_per_apply"],
                              "port_over_reference": ax["oracle_port_over_reference"]}
    return rec


def solver_comparison(nbx, bp, b, r, x, log, tol=1e-8, bp_mg=None, mg_cfg=None, net=None):
    """Time to a 1e-8 relative FP64 residual on the same problem: FP64 Jacobi
    PCG without the iteration cap, and the mixed-precision refinement solvers.
    Each solver runs once untimed, then once timed; the error of each
    solution is reported against the FP64 solution."""
    import torch
    out = {}
    x64 = None

    def run(name, fn):
        nonlocal x64
        for timed in (False, True):
            r.copy_(b)
            x.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = fn()
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
        rel = (st.rdotr / st.rdotr0) ** 0.5 if st.rdotr0 > 0 else 0.0
        rec = {"time_s": round(t, 6), "iterations": st.iterations, "cycles": st.cycles,
               "converged": bool(st.converged), "rel_residual": rel}
        if x64 is None:
            x64 = x.clone()
        else:
            rec["err_vs_fp64"] = float((x - x64).abs().max() / x64.abs().max())
        out[name] = rec
        log(f"[bench] {name}: {rec}")

    run("fp64_jacobi", lambda: bp.pcg(r, x, tol, 20000))
    run("mixed_f32_jacobi", lambda: bp.pcg_mixed(r, x, tol, 20000, prec=nbx.NBX_F32, inner_tol=1e-5, max_cycles=20))
    run("mixed_bf16_jacobi", lambda: bp.pcg_mixed(r, x, tol, 20000, prec=nbx.NBX_BF16, inner_tol=2e-2,
                                                  max_cycles=40))
    if bp_mg is not None:
        orders, degree, ci = mg_cfg
        for name, prec in (("fp64_mg", nbx.NBX_F64), ("mixed_mg_f32cycle", nbx.NBX_F32),
                           ("mixed_mg_bf16cycle", nbx.NBX_BF16), ("mixed_mg_bf16mfma_cycle", nbx.NBX_BF16_MFMA)):
            bp_mg.mg_setup(orders, degree=degree, coarse_iters=ci, prec=prec)
            run(name, lambda: bp_mg.pcg(r, x, tol, 500))
        if net is not None:
            # the network trained on the 16^3 mesh's multigrid pairs, attached
            # to this mesh (its features are this mesh's geometry)
            bp_mg.mg_setup(orders, degree=degree, coarse_iters=ci, prec=nbx.NBX_BF16)
            net.attach(bp_mg)
            bp_mg.mg_set_learned(True)
            run("learned_mixed_mg_bf16cycle", lambda: bp_mg.pcg(r, x, tol, 500))
            bp_mg.mg_set_learned(False)
            out["learned_mixed_mg_bf16cycle"]["gain_vs_same_mg"] = round(
                out["mixed_mg_bf16cycle"]["time_s"] / out["learned_mixed_mg_bf16cycle"]["time_s"], 3)
            out["learned_mixed_mg_bf16cycle"]["inference"] = "bf16 matrix cores"
    set_speedups(out)
    return out


def set_speedups(out):
    """speedup against FP64 Jacobi PCG; null for a solve that stopped at its
    iteration cap without reaching the tolerance (or when FP64 Jacobi did)."""
    base = out["fp64_jacobi"]
    for v in out.values():
        ok = v["converged"] and base["converged"]
        v["speedup_vs_fp64_jacobi"] = round(base["time_s"] / v["time_s"], 3) if ok else None


def flat_time_to_tol(res, prefix):
    """Flat scalar summary of a time-to-tolerance table (kept by parsers that
    drop nested objects): FP64 Jacobi, the fastest converged solver, and the
    learned ones when present."""
    out = {}
    base = res.get("fp64_jacobi")
    if not base:
        return out
    out[f"{prefix}_fp64_jacobi_s"] = base["time_s"] if base["converged"] else None
    done = {k: v for k, v in res.items() if v["converged"] and k != "fp64_jacobi"}
    if done:
        best = min(done, key=lambda k: done[k]["time_s"])
        out[f"{prefix}_best_solver"] = best
        out[f"{prefix}_best_s"] = done[best]["time_s"]
        out[f"{prefix}_best_speedup"] = done[best]["speedup_vs_fp64_jacobi"]
    for k, v in res.items():
        if k.startswith("learned"):
            out[f"{prefix}_{k}_s"] = v["time_s"] if v["converged"] else None
            out[f"{prefix}_{k}_speedup"] = v["speedup_vs_fp64_jacobi"]
            if "gain_vs_same_mg" in v:
                out[f"{prefix}_{k}_gain_vs_same_mg"] = v["gain_vs_same_mg"]
    return out


def field_errors(x, x64, Nf, offset, nlocal):
    xs, rs = x.view(-1)[:Nf * offset].view(Nf, offset)[:, :nlocal], x64.view(-1)[:Nf * offset].view(Nf, offset)[:, :nlocal]
    return [float((xs[f] - rs[f]).abs().max() / rs[f].abs().max()) for f in range(Nf)]


def timed(fn, sync):
    sync()
    t0 = time.perf_counter()
    st = fn()
    sync()
    return st, time.perf_counter() - t0


def run_bp6_n9_24cube(nbx, torch, args):
    """BP6 vector Helmholtz, 3 fields, lambda0 = 1, lambda1 = 0.1 (axhelm_n3),
    N = 9, 24^3 elements, perturbed Dirichlet box; FP64 flexible Jacobi PCG,
    the FP32 / bf16-geometry iterative-refinement solvers and the multigrid
    PCG (orders 9, 4, 1; FP64 and FP32 cycles) to 1e-10; error of each field
    against the FP64 solution."""
    box = nbx.Box.make(9, 24, 24, 24, perturb=0.1, seed=1)
    t0 = time.perf_counter()
    bp = nbx.BP(box, Nfields=3, bpmode=False, lambda1=0.1, precon=nbx.PRECON_JACOBI, flexible=True, fused=True)
    log(f"[bp6] setup {time.perf_counter() - t0:.1f} s E={bp.Nelements} Np={bp.Np}")
    n = 3 * bp.fieldOffset
    b = bp.device_array("rhs", n).clone()
    r, x = torch.empty_like(b), torch.empty_like(b)
    sync = torch.cuda.synchronize
    out = {}
    x64 = None
    runs = [("fp64_jacobi", lambda: bp.pcg(r, x, 1e-10, 20000)),
            ("mixed_f32_jacobi", lambda: bp.pcg_mixed(r, x, 1e-10, 20000, prec=nbx.NBX_F32, inner_tol=1e-5,
                                                      max_cycles=30)),
            ("mixed_bf16_jacobi", lambda: bp.pcg_mixed(r, x, 1e-10, 20000, prec=nbx.NBX_BF16, inner_tol=2e-2,
                                                       max_cycles=80))]
    bm = nbx.BP(box, Nfields=3, bpmode=False, lambda1=0.1, precon=nbx.PRECON_MG)

    def mg(prec):
        def fn():
            return bm.pcg(r, x, 1e-10, 500)
        return fn
    runs += [("fp64_mg", mg(nbx.NBX_F64)), ("mixed_mg_f32cycle", mg(nbx.NBX_F32))]
    mg_prec = {"fp64_mg": nbx.NBX_F64, "mixed_mg_f32cycle": nbx.NBX_F32}
    for name, fn in runs:
        if name in mg_prec:
            bm.mg_setup([9, 4, 1], degree=3, coarse_iters=20, prec=mg_prec[name])
        for _ in range(2):  # the second run is timed
            r.copy_(b)
            x.zero_()
            st, t = timed(fn, sync)
        rec = {"time_s": round(t, 5), "iterations": st.iterations, "cycles": st.cycles,
               "converged": bool(st.converged), "rel_residual": (st.rdotr / st.rdotr0) ** 0.5}
        if x64 is None:
            x64 = x.clone()
        else:
            rec["err_vs_fp64_per_field"] = field_errors(x, x64, 3, bp.fieldOffset, bp.Nlocal)
        out[name] = rec
        log(f"[bp6] {name}: {rec}")
    set_speedups(out)
    ops = bp6_operator_times(nbx, torch, bp)
    return {"config": "BP6 vector Helmholtz N=9, 24^3 elements, 3 fields, lambda1=0.1, tol 1e-10",
            "elements": bp.Nelements, "solvers": out, "operator": ops}


def bp6_operator_times(nbx, torch, bp, sets=2, reps=10):
    """The three-field N=9 Helmholtz operator (axhelm_n3, non-BK) alone on the
    BP6 mesh, per precision: FP64 and FP32 on the vector ALUs, bf16 and fp16
    operands on the matrix cores (k_axhelm_mma16). Each precision gets its own
    `sets` copies of every array (beyond the 256 MB Infinity Cache for every
    precision but bf16 geometry) and is timed over reps x sets launches with
    HIP events on the launch stream; GB/s by axhelm/main.cpp:162-163's count
    (FP64 words: (3*2 + 7 + 2) * Np per element), whatever the precision."""
    E, Np, off = bp.Nelements, bp.Np, bp.fieldOffset
    g64 = bp.device_array("ggeo", E * 7 * Np)
    D64 = bp.device_array("D", 100)
    lam64 = bp.device_array("lambda", 2 * off)
    q64 = torch.rand(3 * off, dtype=torch.float64, device="cuda") * 2 - 1
    ref_bytes = E * Np * (3 * 2 + 7 + 2) * 8
    bf = lambda a: a.float().to(torch.bfloat16).view(torch.int16)
    forms = {"fp64": (nbx.NBX_F64, lambda: (g64.clone(), D64, lam64.clone(), q64.clone())),
             "fp32": (nbx.NBX_F32, lambda: (g64.float(), D64.float(), lam64.float(), q64.float())),
             "bf16_mfma": (nbx.NBX_BF16_MFMA, lambda: (bf(g64), D64.float(), bf(lam64), q64.float())),
             "fp16_mfma": (nbx.NBX_F16, lambda: (g64.float(), D64.float(), lam64.float(), q64.float()))}
    s = torch.cuda.current_stream()
    out = {}
    for name, (prec, make) in forms.items():
        arrs = []
        for _ in range(sets):
            g, D, lam, q = make()
            arrs.append((g, D, lam, q, torch.empty_like(q)))
        for g, D, lam, q, A in arrs:
            nbx.axhelm(9, 3, 0, E, off, g, D, lam, q, A, prec=prec, stream=s)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(s)
        for _ in range(reps):
            for g, D, lam, q, A in arrs:
                nbx.axhelm(9, 3, 0, E, off, g, D, lam, q, A, prec=prec, stream=s)
        ev1.record(s)
        ev1.synchronize()
        t = ev0.elapsed_time(ev1) / 1e3 / (reps * sets)
        out[name] = {"us_per_launch": round(t * 1e6, 1), "GB_s_reference_count": round(ref_bytes / t / 1e9, 1)}
        log(f"[bp6] operator {name}: {out[name]}")
        del arrs
    out["reference_bytes_per_launch"] = ref_bytes
    return out


def train_net(nbx, torch, bp, nseeds, snapshots, steps, lr, log, seed0=1000, mode=1, select_seed=None,
              chunk=250):
    """Pairs from seeded FP64 Jacobi PCG solves, then energy-norm training in
    chunks; with select_seed, the parameters after the chunk whose learned
    PCG needs the fewest iterations on that (held-out) right-hand side are
    kept (model selection on a validation solve)."""
    net = nbx.Net(seed=1)
    net.attach(bp)
    t0 = time.perf_counter()
    k = net.generate_pairs(bp, nseeds=nseeds, seed0=seed0, snapshots=snapshots, tol_exact=1e-11)
    t_pairs = time.perf_counter() - t0
    hist, sel = [], []
    best = None
    if select_seed is not None:
        n = bp.fieldOffset
        vb = torch.empty(n, dtype=torch.float64, device="cuda")
        bp.make_rhs(select_seed, vb)
        vr, vx = torch.empty_like(vb), torch.empty_like(vb)
        bp.set_fallback(10 ** 6, 0.9)
    t0 = time.perf_counter()
    done = 0

    def select():
        nonlocal best
        vr.copy_(vb)
        vx.zero_()
        try:
            st = bp.pcg(vr, vx, 1e-8, 3000)
            its = st.iterations if st.converged else 10 ** 9
        except nbx.NbxError:  # a breakdown of the trial solve rejects these parameters
            its = 10 ** 9
        sel.append((done, its))
        if best is None or its < best[1]:
            best = (done, its, net.params())

    if select_seed is not None:
        select()  # the untrained network (zero output layer: Jacobi) is a candidate too
    while done < steps:
        m = min(chunk, steps - done)
        hist.extend(net.train_energy(bp, steps=m, lr=lr, seed=7 + done, mode=mode).tolist())
        done += m
        if select_seed is not None:
            select()
    if best is not None:
        net.set_params(best[2])
        bp.set_fallback(20, 0.9)
    t_train = time.perf_counter() - t0
    import numpy as np
    hist = np.array(hist)
    log(f"[bench] learned: {k} pairs in {t_pairs:.1f} s, {steps} energy steps (mode {mode}) in {t_train:.1f} s, "
        f"loss {hist[:20].mean():.4f} -> {hist[-100:].mean():.4f}; selection {sel}")
    return net, {"pairs": k, "pair_seeds": [seed0, seed0 + nseeds - 1], "snapshots": snapshots,
                 "pairs_time_s": round(t_pairs, 2), "train_steps": steps, "train_time_s": round(t_train, 2),
                 "loss": "energy" if mode == 0 else "scale-invariant energy",
                 "loss_first20": float(hist[:20].mean()), "loss_last100": float(hist[-100:].mean()),
                 "selection_seed": select_seed, "selection_iterations": sel,
                 "selected_after_steps": best[0] if best else steps}


SNAPS_2000 = [0, 1, 2, 4, 6, 9, 13, 18, 25, 35, 50, 70, 95, 130, 175, 230, 300, 380, 470, 560]


def run_learned_16cube(nbx, torch, args):
    """Learned element preconditioner on the 16^3, N=7 perturbed box: 2000
    (residual, error) pairs from 100 seeded FP64 Jacobi PCG solves (seeds
    1000..1099, 20 snapshots each), energy-norm training with model selection
    on the solve of seed 8999, then test on right-hand sides of unseen seeds
    (9000..9004): iterations and time of the learned flexible PCG (with and
    without the stall fallback) against FP64 Jacobi PCG to 1e-8."""
    box = nbx.Box.make(7, 16, 16, 16, perturb=0.1, seed=1)
    bp = nbx.BP(box, bpmode=True, precon=nbx.PRECON_LEARNED)
    bj = nbx.BP(box, bpmode=True, precon=nbx.PRECON_JACOBI, flexible=True, fused=True)
    net, info = train_net(nbx, torch, bp, 100, SNAPS_2000, args.train_steps, 1e-3, log, mode=args.learn_mode,
                          select_seed=8999)
    n = bp.fieldOffset
    b, r, x = (torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3))
    sync = torch.cuda.synchronize
    rows = []
    for seed in range(9000, 9005):
        bp.make_rhs(seed, b)
        rec = {"seed": seed}
        for name, h, fb, mm in (("jacobi", bj, None, False), ("learned", bp, (20, 0.9), False),
                                ("learned_nofallback", bp, (10 ** 6, 0.9), False),
                                ("learned_mfma_inference", bp, (20, 0.9), True)):
            if fb:
                h.set_fallback(*fb)
            if h is bp:
                net.set_inference(mm)
            for _ in range(2):
                r.copy_(b)
                x.zero_()
                st, t = timed(lambda: h.pcg(r, x, 1e-8, 5000), sync)
            rec[name] = {"iterations": st.iterations, "time_s": round(t, 5), "converged": bool(st.converged),
                         "fallback_iteration": st.fallback_iteration}
        rows.append(rec)
        log(f"[learned] {rec}")
    def avg(k, f):
        return float(sum(row[k][f] for row in rows) / len(rows))
    summary = {k: {"iterations": avg(k, "iterations"), "time_s": avg(k, "time_s")}
               for k in ("jacobi", "learned", "learned_nofallback", "learned_mfma_inference")}
    return {"config": "learned element preconditioner, 16^3 elements N=7, 2000 pairs, validation seeds 9000-9004",
            "training": info, "validation": rows, "mean": summary}


def run_graded_64x64x32(nbx, torch, args):
    """Stress case: 64x64x32 elements N=7 (~45M unknowns), z graded to
    element aspect 1:10 at the bottom, +-0.1 perturbation, tol 1e-10: FP64
    Jacobi PCG, the multigrid PCG with bf16-geometry V-cycle, and the learned
    preconditioner together with the FP32 operator (iterative refinement; the
    network trained on a 16x16x8 copy of the graded mesh)."""
    box = nbx.Box.make(7, 64, 64, 32, perturb=0.1, zaspect=10.0, seed=1)
    sync = torch.cuda.synchronize
    t0 = time.perf_counter()
    bj = nbx.BP(box, bpmode=True, precon=nbx.PRECON_JACOBI, fused=True)
    log(f"[graded] setup {time.perf_counter() - t0:.1f} s E={bj.Nelements}")
    n = bj.fieldOffset
    b = bj.device_array("rhs", n).clone()
    r, x = torch.empty_like(b), torch.empty_like(b)
    out = {}

    def run(name, fn):
        r.copy_(b)
        x.zero_()
        st, t = timed(fn, sync)
        rec = {"time_s": round(t, 4), "iterations": st.iterations, "cycles": st.cycles,
               "converged": bool(st.converged), "rel_residual": (st.rdotr / st.rdotr0) ** 0.5}
        if "fp64_jacobi" in out:
            rec["err_vs_fp64"] = float((x - x64[0]).abs().max() / x64[0].abs().max())
        out[name] = rec
        log(f"[graded] {name}: {rec}")

    x64 = []
    run("fp64_jacobi", lambda: bj.pcg(r, x, 1e-10, args.maxit_graded))
    x64.append(x.clone())
    bj.free()
    bm = nbx.BP(box, bpmode=True, precon=nbx.PRECON_MG)
    bm.mg_setup([7, 3, 1], degree=3, coarse_iters=30, prec=nbx.NBX_BF16)
    run("mixed_mg_bf16cycle", lambda: bm.pcg(r, x, 1e-10, 500))
    if not args.no_learned:
        # the learned step, trained on the multigrid pairs of a 16x16x32 patch with the
        # same element shapes (x, y extent 1/4, the same z grading), then run on this mesh
        from nekbench_amd.learned import train_mg_network
        patch = nbx.BP(nbx.Box.make(7, 16, 16, 32, perturb=0.1, zaspect=10.0, seed=2,
                                    extent=(0.0, 0.25, 0.0, 0.25, 0.0, 1.0)), bpmode=True, precon=nbx.PRECON_MG)
        patch.mg_setup([7, 3, 1], degree=3, coarse_iters=30, prec=nbx.NBX_BF16)
        mnet, minfo = train_mg_network(nbx, torch, patch, seeds=(1000, args.mg_pair_seeds), kcount=8,
                                       steps=args.train_steps_mg, alpha=args.mg_alpha, log=log)
        patch.free()
        mnet.set_inference(True)
        mnet.attach(bm)
        bm.mg_set_learned(True)
        run("learned_mixed_mg_bf16cycle", lambda: bm.pcg(r, x, 1e-10, 500))
        out["learned_mixed_mg_bf16cycle"]["training"] = dict(minfo, mesh="16x16x32 patch, extent 1/4 x 1/4 x 1")
        out["learned_mixed_mg_bf16cycle"]["gain_vs_same_mg"] = round(
            out["mixed_mg_bf16cycle"]["time_s"] / out["learned_mixed_mg_bf16cycle"]["time_s"], 3)
        bm.mg_set_learned(False)
        mnet.free()
    bm.free()
    small = nbx.BP(nbx.Box.make(7, 16, 16, 8, perturb=0.1, zaspect=10.0, seed=2), bpmode=True,
                   precon=nbx.PRECON_LEARNED)
    net, info = train_net(nbx, torch, small, 30, SNAPS_2000[:14], args.train_steps // 2, 1e-3, log,
                          mode=args.learn_mode, select_seed=8999)
    bl = nbx.BP(box, bpmode=True, precon=nbx.PRECON_LEARNED)
    net.attach(bl)
    run("learned_mixed_f32", lambda: bl.pcg_mixed(r, x, 1e-10, args.maxit_graded, prec=nbx.NBX_F32, inner_tol=1e-5,
                                                  max_cycles=20))
    set_speedups(out)
    return {"config": "64x64x32 elements N=7, z-graded aspect 1:10, perturbed, tol 1e-10",
            "elements": 64 * 64 * 32, "learned_training": info, "solvers": out}


def run_learned_mg(nbx, torch, args):
    """Learned multigrid: the element network as the last step of the fine
    level of every V-cycle (bf16-geometry cycle, orders 7, 3, 1), trained on
    the mesh's own multigrid pairs (nekbench_amd.learned), against the same
    multigrid without the network, on a uniform 16^3 mesh (tol 1e-8) and a
    z-graded 16x16x8 mesh (aspect 1:10, tol 1e-10): iterations and time to
    tolerance on five unseen right-hand sides.  --mg-variants lists training
    variants kfirst:kcount:mode:lr:alpha (pairs of iterates kfirst..,
    loss mode 0 energy / 1 scale-invariant, Adam rate, damped-Jacobi start)."""
    import numpy as np
    from nekbench_amd.learned import train_mg_network
    sync = torch.cuda.synchronize
    variants = []
    for v in args.mg_variants.split(","):
        kf, kc, mode, lr, alpha = v.split(":")
        variants.append((v, int(kf), int(kc), int(mode), float(lr), float(alpha)))
    cases = [("uniform_16cube", nbx.Box.make(7, 16, 16, 16, perturb=0.1, seed=1), 1e-8, 8),
             ("graded_16x16x8", nbx.Box.make(7, 16, 16, 8, perturb=0.1, zaspect=10.0, seed=2), 1e-10, 8)]
    out = {}
    for name, box, tol, ci in cases:
        bm = nbx.BP(box, bpmode=True, precon=nbx.PRECON_MG)
        bm.mg_setup([7, 3, 1], degree=args.mg_degree, coarse_iters=ci, prec=nbx.NBX_BF16)
        n = bm.fieldOffset
        b, r, x = (torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3))
        seeds = list(range(9000, 9000 + args.mg_eval_seeds))

        def evaluate():
            its, ts = [], []
            for seed in seeds:
                bm.make_rhs(seed, b)
                for _ in range(2):
                    r.copy_(b)
                    x.zero_()
                    st, t = timed(lambda: bm.pcg(r, x, tol, 500), sync)
                its.append(st.iterations if st.converged else None)
                ts.append(round(t, 5))
            ok = all(i is not None for i in its)
            return {"iterations": its, "time_s": ts, "mean_iterations": float(np.mean(its)) if ok else None,
                    "mean_time_s": float(np.mean(ts)), "converged": ok}

        res = {"tol": tol, "coarse_iterations": ci}
        # the plain cycle with one more Chebyshev step: the learned step costs
        # about one fine-level smoothing step, so it must beat that too
        bm.mg_setup([7, 3, 1], degree=args.mg_degree + 1, coarse_iters=ci, prec=nbx.NBX_BF16)
        res[f"mg_degree{args.mg_degree + 1}"] = evaluate()
        log(f"[learned-mg] {name} plain multigrid, degree {args.mg_degree + 1}: {res[f'mg_degree{args.mg_degree + 1}']}")
        bm.mg_setup([7, 3, 1], degree=args.mg_degree, coarse_iters=ci, prec=nbx.NBX_BF16)
        res["mg"] = evaluate()
        log(f"[learned-mg] {name} plain multigrid {res['mg']}")
        for tag, kf, kc, mode, lr, alpha in variants:
            net, info = train_mg_network(nbx, torch, bm, seeds=(1000, args.mg_pair_seeds), kfirst=kf, kcount=kc,
                                         steps=args.train_steps_mg, lr=lr, alpha=alpha, mode=mode, log=log)
            rec = {"training": info}
            for inf, mfma in (("fp32_inference", False), ("mfma_inference", True)):
                net.set_inference(mfma)
                rec[inf] = evaluate()
            base = res["mg"]["mean_time_s"]
            rec["gain_vs_same_mg"] = round(base / rec["mfma_inference"]["mean_time_s"], 3)
            rec[f"gain_vs_mg_degree{args.mg_degree + 1}"] = round(
                res[f"mg_degree{args.mg_degree + 1}"]["mean_time_s"] / rec["mfma_inference"]["mean_time_s"], 3)
            log(f"[learned-mg] {name} variant {tag}: fp32 {rec['fp32_inference']} mfma {rec['mfma_inference']}")
            res[f"learned_{tag}"] = rec
            bm.mg_set_learned(False)
            net.free()
        out[name] = res
        bm.free()
    return {"config": "learned multigrid step vs the same multigrid, bf16-geometry cycle (7,3,1), degree "
                      f"{args.mg_degree}", "cases": out}


def run_sustained_32cube(nbx, torch, args):
    """Per-launch durations of the stand-alone FP64 operator (32^3, N=7, four rotating 1.2 GB buffer
    sets) back to back and with 2 ms idle before each launch, and of a triad moving 1.5 GB per launch
    back to back: separates the kernel's own rate from the board's response to sustained load."""
    N = 7
    box = nbx.Box.make(N, 32, 32, 32, perturb=0.1, seed=1)
    bp = nbx.BP(box, bpmode=True, precon=nbx.PRECON_JACOBI)
    E, n = bp.Nelements, bp.fieldOffset
    ggeo = bp.device_array("ggeo", E * 7 * bp.Np)
    D = bp.device_array("D", (N + 1) ** 2)
    x = torch.rand(n, dtype=torch.float64, device="cuda:0")
    sets = [(ggeo, x, torch.empty_like(x))] + [(ggeo.clone(), torch.rand_like(x), torch.empty_like(x))
                                               for _ in range(ROOF_SETS - 1)]
    s = torch.cuda.current_stream()
    count = 40

    def timed(launch, k, idle):
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(k + 1)]
        out = []
        if idle:
            for i in range(k):
                s.synchronize()
                time.sleep(idle)
                evs[0].record(s)
                launch(i)
                evs[1].record(s)
                evs[1].synchronize()
                out.append(evs[0].elapsed_time(evs[1]) * 1e3)
            return out
        s.synchronize()
        evs[0].record(s)
        for i in range(k):
            launch(i)
            evs[i + 1].record(s)
        evs[k].synchronize()
        return [evs[i].elapsed_time(evs[i + 1]) * 1e3 for i in range(k)]

    def ax(i):
        g_, q_, a_ = sets[i % ROOF_SETS]
        nbx.axhelm(N, 1, 1, E, n, g_, D, None, q_, a_, stream=s)

    res = {"workload": "FP64 BK axhelm N=7 32^3 (1.208 GB per launch, 4 rotating sets); triad 3 x 512 MB"}
    for _ in range(ROOF_SETS):
        ax(_)
    res["axhelm_back_to_back_us"] = [round(v, 1) for v in timed(ax, count, 0.0)]
    time.sleep(1.0)
    res["axhelm_idle_2ms_us"] = [round(v, 1) for v in timed(ax, count, 0.002)]
    del sets
    m = 1 << 26
    tri = [(torch.rand(m, dtype=torch.float64, device="cuda:0"), torch.rand(m, dtype=torch.float64, device="cuda:0"),
            torch.empty(m, dtype=torch.float64, device="cuda:0")) for _ in range(2)]

    def triad(i):
        a, b_, c = tri[i % 2]
        nbx.lib().nbx_triad(m, 0.5, a.data_ptr(), b_.data_ptr(), c.data_ptr(), s.cuda_stream)

    time.sleep(1.0)
    res["triad_back_to_back_us"] = [round(v, 1) for v in timed(triad, 60, 0.0)]
    for k in ("axhelm_back_to_back_us", "axhelm_idle_2ms_us", "triad_back_to_back_us"):
        v = res[k]
        log(f"[sustained] {k}: first 10 mean {sum(v[:10]) / 10:.1f}, last 10 mean {sum(v[-10:]) / 10:.1f}, "
            f"min {min(v):.1f}, max {max(v):.1f}")
    return res


def mesh16_measurements(nbx, torch, args, log, axhelm_only=False):
    """The 16^3-element mesh (N=7, perturbed, seed 1): FP64 BK axhelm
    with 5 rotating buffer sets of 151 MB each (755 MB, beyond the 256 MB
    Infinity Cache), HIP events on the launch stream; and time to a 1e-8
    relative residual of FP64 Jacobi PCG against the multigrid solvers on the
    Dirichlet BP5 problem of the same mesh."""
    box = nbx.Box.make(7, 16, 16, 16, perturb=0.1, seed=1)
    bj = nbx.BP(box, bpmode=True, precon=nbx.PRECON_JACOBI, fused=True)
    E, Np, n = bj.Nelements, bj.Np, bj.fieldOffset
    g0 = bj.device_array("ggeo", E * 7 * Np)
    D = bj.device_array("D", 64)
    sets = []
    for k in range(5):
        sets.append((g0.clone(), torch.rand(n, dtype=torch.float64, device="cuda") * 2 - 1,
                     torch.empty(n, dtype=torch.float64, device="cuda")))
    s = torch.cuda.current_stream()
    for g, q, a in sets:
        nbx.axhelm(7, 1, 1, E, n, g, D, None, q, a, stream=s)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    ev0.record(s)
    for _ in range(reps):
        for g, q, a in sets:
            nbx.axhelm(7, 1, 1, E, n, g, D, None, q, a, stream=s)
    ev1.record(s)
    ev1.synchronize()
    t_ax = ev0.elapsed_time(ev1) / 1e3 / (reps * len(sets))
    bytes_ = E * Np * 9 * 8
    out = {"axhelm_fp64_us": round(t_ax * 1e6, 2), "axhelm_GB_s": round(bytes_ / t_ax / 1e9, 1),
           "axhelm_frac_of_8TBs": round(bytes_ / t_ax / 8e12, 4), "axhelm_gdofs": round(E * 343 / t_ax / 1e9, 2),
           "buffer_sets": len(sets), "bytes_per_apply": bytes_}
    del sets
    # the practical ceiling at this size: a streaming triad c = a + alpha b
    # (nbx_triad, the bw benchmark's kernel) moving the same bytes per launch,
    # rotated over the same number of buffer sets, timed the same way
    nt = bytes_ // 24
    tri = [tuple(torch.rand(nt, dtype=torch.float64, device="cuda") for _ in range(3)) for _ in range(5)]
    for a, b_, c in tri:
        nbx.lib().nbx_triad(nt, 0.5, a.data_ptr(), b_.data_ptr(), c.data_ptr(), s.cuda_stream)
    ev0.record(s)
    for _ in range(reps):
        for a, b_, c in tri:
            nbx.lib().nbx_triad(nt, 0.5, a.data_ptr(), b_.data_ptr(), c.data_ptr(), s.cuda_stream)
    ev1.record(s)
    ev1.synchronize()
    t_tri = ev0.elapsed_time(ev1) / 1e3 / (reps * len(tri))
    del tri
    out.update({"triad_same_bytes_us": round(t_tri * 1e6, 2), "triad_same_bytes_GB_s": round(24 * nt / t_tri / 1e9, 1),
                "axhelm_over_triad_time": round(t_ax / t_tri, 3)})
    log(f"[bench] 16^3 axhelm: {out}")
    if axhelm_only:
        return out, None
    b = bj.device_array("rhs", n).clone()
    r, x = torch.empty_like(b), torch.empty_like(b)
    sync = torch.cuda.synchronize
    res = {}

    def run(name, h):
        for _ in range(2):
            r.copy_(b)
            x.zero_()
            st, t = timed(lambda: h.pcg(r, x, 1e-8, 20000), sync)
        res[name] = {"iterations": st.iterations, "time_s": round(t, 5), "converged": bool(st.converged)}

    run("fp64_jacobi", bj)
    bm = nbx.BP(box, bpmode=True, precon=nbx.PRECON_MG)
    for name, prec in (("fp64_mg", nbx.NBX_F64), ("mixed_mg_f32cycle", nbx.NBX_F32),
                       ("mixed_mg_bf16cycle", nbx.NBX_BF16)):
        # the smaller coarse problem (15^3 unknowns) needs fewer coarse
        # iterations than 32^3: 10 were the fastest of 5, 6, 7, 8, 10 (bf16 cycle)
        bm.mg_setup([7, 3, 1], degree=3, coarse_iters=10, prec=prec)
        run(name, bm)
    # learned multigrid: the bf16-geometry cycle plus the network's step on
    # the fine level, trained here on this mesh's multigrid pairs
    net = None
    if not args.no_learned:
        from nekbench_amd.learned import train_mg_network
        net, info = train_mg_network(nbx, torch, bm, seeds=(1000, args.mg_pair_seeds), kcount=8,
                                     steps=args.train_steps_mg, alpha=args.mg_alpha, log=log)
        net.set_inference(True)
        run("learned_mixed_mg_bf16cycle", bm)
        bm.mg_set_learned(False)
        res["learned_mixed_mg_bf16cycle"]["gain_vs_same_mg"] = round(
            res["mixed_mg_bf16cycle"]["time_s"] / res["learned_mixed_mg_bf16cycle"]["time_s"], 3)
        res["learned_mixed_mg_bf16cycle"]["inference"] = "bf16 matrix cores"
        out["learned_mg_training"] = info
    set_speedups(res)
    out["time_to_1e-8"] = res
    log(f"[bench] 16^3 time to 1e-8: {res}")
    return out, net


def profile_mg_pass(nbx, torch, NX):
    import numpy as np
    box = nbx.Box.make(7, NX, NX, NX, perturb=0.1, seed=1)
    bp = nbx.BP(box, bpmode=True, precon=nbx.PRECON_MG)
    bp.mg_setup([7, 3, 1], degree=3, coarse_iters=20, prec=nbx.NBX_BF16)
    n = bp.fieldOffset
    b = bp.device_array("rhs", n).clone()
    r, x = torch.empty_like(b), torch.empty_like(b)
    for _ in range(2):
        r.copy_(b)
        x.zero_()
        st = bp.pcg(r, x, 1e-8, 500)
    log(f"[bench] multigrid bf16 cycle: {st.iterations} iterations, {st.elapsed_s * 1e3:.1f} ms")
    bl = nbx.BP(box, bpmode=True, precon=nbx.PRECON_LEARNED)
    net = nbx.Net(seed=3)
    W = net.params()
    W[2176:3040] = np.random.default_rng(0).uniform(-0.05, 0.05, 864).astype(np.float32)  # nonzero output layer
    net.set_params(W)
    net.attach(bl)
    bl.update_jacobi()
    zl = torch.empty_like(b)
    for _ in range(5):
        net.apply(bl, b, zl)
    net.set_inference(True)  # the matrix-core inference
    for _ in range(5):
        net.apply(bl, b, zl)
    torch.cuda.synchronize()
    # the learned multigrid step (its damped-Jacobi start network, matrix-core inference)
    from nekbench_amd.learned import damped_jacobi_params
    net2 = nbx.Net(seed=4)
    net2.set_params(damped_jacobi_params(net2.params(), 0.5))
    net2.set_inference(True)
    net2.attach(bp)
    bp.mg_set_learned(True)
    for _ in range(2):
        r.copy_(b)
        x.zero_()
        st = bp.pcg(r, x, 1e-8, 500)
    log(f"[bench] learned multigrid (damped-Jacobi start): {st.iterations} iterations, {st.elapsed_s * 1e3:.1f} ms")
    log("[bench] profile pass done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=32)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--maxit", type=int, default=500)
    ap.add_argument("--ax-launches", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the time-to-1e-8 comparison of the FP64, mixed and learned solvers")
    ap.add_argument("--study", default="pcg32",
                    choices=("pcg32", "axhelm16", "bp6", "learned16", "graded", "learned_mg", "sustained"),
                    help="pcg32: the PCG benchmark line on 32^3 (default); axhelm16: the 16^3 FP64 axhelm alone "
                         "(rotating buffer sets); bp6: BP6 N=9 24^3 mixed precision and its operator per "
                         "precision; sustained: per-launch times of the 32^3 FP64 axhelm back to back and "
                         "with idle gaps, and of a triad; learned16: learned "
                         "preconditioner on 16^3; graded: 64x64x32 graded stress case (all but pcg32 print "
                         "one JSON line of their own)")
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--train-steps-mg", type=int, default=3000,
                    help="Adam steps of the learned multigrid step's network (trained on the 16^3 mesh)")
    ap.add_argument("--mg-pair-seeds", type=int, default=100,
                    help="right-hand sides of the learned multigrid's training pairs (8 pairs each)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned multigrid solver")
    ap.add_argument("--mg-degree", type=int, default=3, help="Chebyshev degree of the learned-multigrid study")
    ap.add_argument("--mg-variants", default="0:8:1:3e-3:0.3,0:8:0:1e-3:0.5",
                    help="learned-multigrid study: comma list of kfirst:kcount:mode:lr:alpha training variants")
    ap.add_argument("--mg-eval-seeds", type=int, default=5, help="right-hand sides of the learned-multigrid study")
    ap.add_argument("--mg-alpha", type=float, default=0.3,
                    help="damping of the learned step's damped-Jacobi start (1 - alpha multiplies Jacobi)")
    ap.add_argument("--learn-mode", type=int, default=1, choices=(0, 1),
                    help="energy-norm loss: 0 fixed scale, 1 scale-invariant")
    ap.add_argument("--maxit-graded", type=int, default=30000)
    ap.add_argument("--profile-mg", action="store_true",
                    help="short sequence for a rocprofv3 kernel trace: two multigrid PCG solves (bf16-geometry "
                         "V-cycle, degree 3, 20 coarse iterations) to 1e-8, five network applies with the FP32 and "
                         "five with the matrix-core inference, two learned-multigrid solves; prints no JSON")
    ap.add_argument("--profile-pass", action="store_true",
                    help="short fixed sequence for rocprofv3 counter runs: a 1 GiB device copy (calibration), "
                         "8 axhelm launches over the roofline's 4 rotating buffer sets, one 10-iteration PCG "
                         "solve; prints no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the element sharding over several GPUs is a design, not built)", file=sys.stderr)
        sys.exit(2)

    import torch
    import nekbench_amd as nbx

    if args.profile_mg:
        profile_mg_pass(nbx, torch, args.nx)
        return
    if args.study != "pcg32":
        torch.cuda.set_device(0)
        if args.study == "axhelm16":
            res = mesh16_measurements(nbx, torch, args, log, axhelm_only=True)[0]
        else:
            res = {"bp6": run_bp6_n9_24cube, "learned16": run_learned_16cube,
                   "graded": run_graded_64x64x32, "learned_mg": run_learned_mg,
                   "sustained": run_sustained_32cube}[args.study](nbx, torch, args)
        print(json.dumps(res), flush=True)
        return

    N, NX = 7, args.nx
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    box = nbx.Box.make(N, NX, NX, NX, perturb=0.1, seed=1)
    t0 = time.perf_counter()
    bp = nbx.BP(box, bpmode=True, precon=nbx.PRECON_JACOBI)  # Poisson: lambda0=1, lambda1=0
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    log(f"[bench] setup {setup_s:.2f} s: E={bp.Nelements} Nlocal={bp.Nlocal}")
    n = bp.fieldOffset
    b = bp.device_array("rhs", n).clone()
    r = torch.empty_like(b)
    x = torch.empty_like(b)

    def step():
        r.copy_(b)
        x.zero_()
        return bp.pcg(r, x, args.tol, args.maxit)

    if args.profile_pass:
        src = torch.zeros(1 << 27, dtype=torch.float64, device=dev)   # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        ggeo = bp.device_array("ggeo", bp.Nelements * 7 * bp.Np)
        D = bp.device_array("D", (N + 1) ** 2)
        x.copy_(b)
        # the roofline's rotating buffer sets: each launch reads a set the
        # previous ROOF_SETS - 1 launches did not touch
        sets = [(ggeo, x, torch.empty_like(x))] + [(ggeo.clone(), torch.rand_like(x) * 2 - 1, torch.empty_like(x))
                                                   for _ in range(ROOF_SETS - 1)]
        for k in range(2 * ROOF_SETS):
            g_, q_, a_ = sets[k % ROOF_SETS]
            nbx.axhelm(N, 1, 1, bp.Nelements, n, g_, D, None, q_, a_)
        del sets
        r.copy_(b)
        x.zero_()
        bp.pcg(r, x, 0.0, 10)
        torch.cuda.synchronize()
        log("[bench] profile pass done")
        return

    for _ in range(args.warmup):
        st = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    its = st.iterations
    E = bp.Nelements
    gdofs = its * E * N ** 3 / t / 1e9
    log(f"[bench] PCG: {its} iterations, converged={st.converged}, rel res "
        f"{(st.rdotr / st.rdotr0) ** 0.5:.3e}, {t * 1e3:.2f} ms/solve ({t / max(its, 1) * 1e6:.1f} us/iter)")

    # axhelm launches on the solve's mesh with ROOF_SETS rotating (ggeo, q, Aq)
    # buffer sets (4 x 1.2 GB, far beyond the 256 MB Infinity Cache, so no
    # launch finds its inputs in cache), HIP events on the launch stream.
    # Each launch is timed by its own event pair with 2 ms of idle time before
    # it: back to back, after about a dozen launches (2-3 ms) this FP64 kernel
    # slows from ~178 to 200-240 us while a triad of the same bytes does not,
    # i.e. the board's power management lowers the clocks; the back-to-back
    # average is reported beside it.
    ggeo = bp.device_array("ggeo", E * 7 * bp.Np)
    D = bp.device_array("D", (N + 1) ** 2)
    Ap = torch.empty_like(x)
    s = torch.cuda.current_stream()
    roof_sets = [(ggeo, x, Ap)] + [(ggeo.clone(), torch.rand_like(x) * 2 - 1, torch.empty_like(x))
                                   for _ in range(ROOF_SETS - 1)]
    for g_, q_, a_ in roof_sets:
        nbx.axhelm(N, 1, 1, E, n, g_, D, None, q_, a_, stream=s)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(1, args.ax_launches // ROOF_SETS)
    per_launch = []
    for i in range(reps * ROOF_SETS):
        g_, q_, a_ = roof_sets[i % ROOF_SETS]
        s.synchronize()
        time.sleep(0.002)
        ev0.record(s)
        nbx.axhelm(N, 1, 1, E, n, g_, D, None, q_, a_, stream=s)
        ev1.record(s)
        ev1.synchronize()
        per_launch.append(ev0.elapsed_time(ev1) / 1e3)
    ax_s = sum(per_launch) / len(per_launch)
    ev0.record(s)
    for _ in range(reps):
        for g_, q_, a_ in roof_sets:
            nbx.axhelm(N, 1, 1, E, n, g_, D, None, q_, a_, stream=s)
    ev1.record(s)
    ev1.synchronize()
    ax_b2b_s = ev0.elapsed_time(ev1) / 1e3 / (reps * ROOF_SETS)
    del roof_sets
    ax_bytes = E * bp.Np * (2 + 7) * 8
    ax_gbs = ax_bytes / ax_s / 1e9
    log(f"[bench] axhelm BK N=7 E={E}: {ax_s * 1e6:.1f} us/launch over {ROOF_SETS} buffer sets "
        f"(min {min(per_launch) * 1e6:.1f}, max {max(per_launch) * 1e6:.1f}), {ax_gbs:.0f} GB/s, "
        f"{E * N ** 3 / ax_s / 1e9:.1f} GDOF/s; back to back {ax_b2b_s * 1e6:.1f} us/launch")

    # the other precisions of the same operator on the same mesh (FP32 vectors)
    variants = {}
    g32, D32, q32 = ggeo.float(), D.float(), x.float().contiguous()
    g16 = ggeo.to(torch.bfloat16).view(torch.int16)
    A32 = torch.empty_like(q32)
    for name, prec, g, bpn in (("f32", nbx.NBX_F32, g32, 36), ("bf16_geometry", nbx.NBX_BF16, g16, 22),
                               ("bf16_mfma", nbx.NBX_BF16_MFMA, g16, 22), ("f16_mfma", nbx.NBX_F16, g32, 36)):
        for _ in range(3):
            nbx.axhelm(N, 1, 1, E, n, g, D32, None, q32, A32, prec=prec, stream=s)
        ev0.record(s)
        for _ in range(args.ax_launches):
            nbx.axhelm(N, 1, 1, E, n, g, D32, None, q32, A32, prec=prec, stream=s)
        ev1.record(s)
        ev1.synchronize()
        t_v = ev0.elapsed_time(ev1) / 1e3 / args.ax_launches
        variants[name] = {"us_per_launch": round(t_v * 1e6, 2), "bytes_per_node": bpn,
                          "GB_s": round(E * bp.Np * bpn / t_v / 1e9, 1), "gdofs": round(E * N ** 3 / t_v / 1e9, 2)}
    log(f"[bench] axhelm variants: {variants}")
    del g32, g16, q32, A32

    mesh16, net = (None, None) if args.no_compare else mesh16_measurements(nbx, torch, args, log)

    compare = None
    if not args.no_compare:
        t1 = time.perf_counter()
        mg_cfg = ([7, 3, 1], 3, 20)  # orders, Chebyshev degree, coarse PCG iterations
        bp_mg = nbx.BP(box, bpmode=True, precon=nbx.PRECON_MG)
        bp_mg.mg_setup(mg_cfg[0], degree=mg_cfg[1], coarse_iters=mg_cfg[2])
        log(f"[bench] multigrid setup {time.perf_counter() - t1:.1f} s (orders 7,3,1)")
        compare = solver_comparison(nbx, bp, b, r, x, log, bp_mg=bp_mg, mg_cfg=mg_cfg, net=net)
        bp_mg.free()

    traffic = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists() and NX == 32:
        traffic = json.loads(tfile.read_text()).get("axhelm_bk_f64_n7_32cube", {}).get("traffic_bytes_per_launch")
    cpu = None if args.no_cpu else run_cpu_baseline(NX, 0 if args.no_compare else 16, log)
    if mesh16 is not None and cpu and "axhelm" in cpu:
        mesh16["cpu_reference_axhelm"] = dict(cpu["axhelm"], cores=cpu["cores"]["threads"], cpu=cpu["cores"]["cpu"])
    flat = {}
    if compare:
        flat.update(flat_time_to_tol(compare, "t1e8_32cube"))
    if mesh16 and "time_to_1e-8" in mesh16:
        flat.update(flat_time_to_tol(mesh16["time_to_1e-8"], "t1e8_16cube"))
    if mesh16:
        flat.update({"axhelm16_us": mesh16["axhelm_fp64_us"], "axhelm16_frac_of_8TBs": mesh16["axhelm_frac_of_8TBs"],
                     "triad16_same_bytes_us": mesh16.get("triad_same_bytes_us")})
    out = {
        "metric": METRIC, "value": round(gdofs, 4), "unit": "GDOF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"nekBone FP64 Jacobi PCG, {NX}^3 elements N=7 Dirichlet perturbed box, "
                               f"tol {args.tol:g} or {args.maxit} iterations", "elements": E, "N": N,
                   "iterations": its, "converged": bool(st.converged),
                   # wall time of one solve; a time to tolerance only when converged is true
                   "time_s": round(t, 6), "final_rel_residual": (st.rdotr / st.rdotr0) ** 0.5,
                   "axhelm_gdofs": round(E * N ** 3 / ax_s / 1e9, 3),
                   # axhelm/main.cpp:165-168: (12 Nq + 15) flop per node (BK); FP64 vector peak 78.6 TFLOP/s
                   "axhelm_fp64_tflops": round(E * bp.Np * (12 * (N + 1) + 15) / ax_s / 1e12, 2),
                   "axhelm_fp64_frac_of_78p6": round(E * bp.Np * (12 * (N + 1) + 15) / ax_s / 78.6e12, 4),
                   "axhelm_variants": variants,
                   "parallelism": "single GPU", "setup_s": round(setup_s, 3),
                   **flat,
                   "time_to_1e-8": compare, "mesh16": mesh16},
        "roofline": {"bound": "hbm", "achieved": round(ax_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ax_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_axhelm_bk8_split", "bytes_per_launch": ax_bytes,
                     "us_per_launch": round(ax_s * 1e6, 3), "buffer_sets": ROOF_SETS,
                     "timing": "per-launch HIP events, 2 ms idle before each launch",
                     "us_per_launch_back_to_back": round(ax_b2b_s * 1e6, 3),
                     "frac_back_to_back": round(ax_bytes / ax_b2b_s / 1e9 / HBM_PEAK_GBS, 4),
                     "frac_by_counter_bytes": (round(traffic / ax_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None)},
        "cpu_baseline": cpu_baseline_record(cpu),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
