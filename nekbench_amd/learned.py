"""Training of the learned multigrid step (host driver over the C ABI).

The network of nbx_net_* (one element, 8^3 nodes, N = 7) is applied at the
end of every V-cycle to the fine-level residual the cycle leaves behind
(nbx_bp_mg_set_learned).  It is trained on what it will see there:

  pairs      nbx_net_generate_mg_pairs: seeded right-hand sides, the
             multigrid PCG's iterates x_k, and the residual / error one
             V-cycle leaves on them, (r' = r_k - A M r_k, e' = e_k - M r_k);
  loss       nbx_net_train_energy: mode 0, (z - e')^T A (z - e') / e'^T A e'
             with z = mask(gs(w invDiag (r' + s f(r')))), the A-norm of the
             error left after the learned correction relative to before it
             (below 1 exactly when the step reduces the error); mode 1 (the
             default), its scale-invariant form;
  start      damped Jacobi, z = (1 - alpha) invDiag r' (two hidden channels
             carry relu(r^) and relu(-r^) through the network, the output
             layer scales them), so training starts from a step that already
             reduces the error instead of from plain Jacobi;
  selection  after each chunk of Adam steps, the multigrid PCG with the
             learned step runs a fixed number of iterations on a held-out
             right-hand side; the parameters with the smallest final residual
             are kept.

The reference has no learned component; its preconditioner hook is
BPPreconditioner (nekBone/solve.cpp:167-179) inside the flexible PCG
(nekBone/solve.cpp:92-95), which is where the learned multigrid sits.
"""
from __future__ import annotations

import time

import numpy as np

# parameter layout of the element network (nbx_net.h)
OFF_WR, OFF_WG, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 864, 1088, 1120, 2144, 2176, 3040
CENTER_TAP = 13  # (dz, dy, dx) = (0, 0, 0)


def damped_jacobi_params(params: np.ndarray, alpha: float) -> np.ndarray:
    """Parameters whose network output is y = -alpha r^ exactly, so that the
    element output is (1 - alpha) invDiag r: channels 0 and 1 carry relu(r^)
    and relu(-r^) (centre tap +-1, no geometry, zero bias), layer 2 passes
    them through, the output layer's centre tap weighs them -alpha and
    +alpha; every other output weight is zero.  alpha is a bf16 value."""
    P = np.array(params, dtype=np.float32, copy=True)
    for c, sign in ((0, 1.0), (1, -1.0)):
        for tap in range(27):                         # residual taps into channel c: centre only
            P[OFF_WR + tap * 32 + c] = sign if tap == CENTER_TAP else 0.0
        for q in range(7):                            # no geometry into channel c
            P[OFF_WG + q * 32 + c] = 0.0
        P[OFF_B1 + c] = 0.0
        for ci in range(32):                          # column c of W2 (in x out) passes channel c through
            P[OFF_W2 + ci * 32 + c] = 1.0 if ci == c else 0.0
        P[OFF_B2 + c] = 0.0
    P[OFF_W3:OFF_B3] = 0.0
    P[OFF_W3 + CENTER_TAP * 32 + 0] = -alpha
    P[OFF_W3 + CENTER_TAP * 32 + 1] = alpha
    P[OFF_B3] = 0.0
    return P


def validation_residual(bp, b, r, x, iters, torch):
    """Relative weighted residual after `iters` multigrid-PCG iterations on b."""
    r.copy_(b)
    x.zero_()
    st = bp.pcg(r, x, 0.0, iters)
    return (st.rdotr / st.rdotr0) ** 0.5 if st.rdotr0 > 0 else 0.0


def train_mg_network(nbx, torch, bp, *, seeds=(1000, 100), kfirst=0, kcount=8, steps=3000, chunk=250, lr=3e-3,
                     alpha=0.3, mode=1, val_seed=8999, val_iters=6, net_seed=1, log=print):
    """Train a network for the learned step of `bp` (precon MG, multigrid set
    up, N = 7, one field) on the pairs of iterates kfirst .. kfirst+kcount-1
    with the energy loss of nbx_net_train_energy (mode 0: the A-norm error
    after the step relative to before it; 1: its scale-invariant form).
    Returns (net, info); the network is attached to bp and the learned step
    is left on."""
    net = nbx.Net(seed=net_seed)
    net.attach(bp)
    net.set_params(damped_jacobi_params(net.params(), alpha))
    t0 = time.perf_counter()
    npairs = net.generate_mg_pairs(bp, nseeds=seeds[1], seed0=seeds[0], kcount=kcount, kfirst=kfirst,
                                   tol_exact=1e-11)
    t_pairs = time.perf_counter() - t0
    n = bp.fieldOffset
    vb = torch.empty(n, dtype=torch.float64, device="cuda")
    bp.make_rhs(val_seed, vb)
    vr, vx = torch.empty_like(vb), torch.empty_like(vb)
    bp.mg_set_learned(False)
    plain = validation_residual(bp, vb, vr, vx, val_iters, torch)
    bp.mg_set_learned(True)
    sel = []
    best = None

    def select(done):
        nonlocal best
        try:
            v = validation_residual(bp, vb, vr, vx, val_iters, torch)
        except nbx.NbxError:  # a breakdown of the trial solve rejects these parameters
            v = float("inf")
        if not np.isfinite(v):
            v = float("inf")
        sel.append((done, v))
        if best is None or v < best[1]:
            best = (done, v, net.params())

    select(0)
    hist = []
    done = 0
    t0 = time.perf_counter()
    while done < steps:
        m = min(chunk, steps - done)
        hist.extend(net.train_energy(bp, steps=m, lr=lr, seed=7 + done, mode=mode).tolist())
        done += m
        select(done)
    net.set_params(best[2])
    t_train = time.perf_counter() - t0
    hist = np.array(hist) if hist else np.zeros(1)
    info = {"pairs": npairs, "pair_seeds": [seeds[0], seeds[0] + seeds[1] - 1],
            "iterates": [kfirst, kfirst + kcount - 1], "pairs_time_s": round(t_pairs, 2), "train_steps": steps,
            "train_time_s": round(t_train, 2), "lr": lr,
            "loss": ("energy (A-norm error after the step / before)" if mode == 0 else
                     "scale-invariant energy"), "init": f"damped Jacobi, alpha {alpha}",
            "loss_first20": float(hist[:20].mean()), "loss_last100": float(hist[-100:].mean()),
            "validation_seed": val_seed, "validation_iterations": val_iters,
            "validation_rel_residual_plain_mg": plain,
            "validation_rel_residual": [(d, float(v)) for d, v in sel], "selected_after_steps": best[0]}
    log(f"[learned-mg] {npairs} pairs in {t_pairs:.1f} s; {steps} steps in {t_train:.1f} s; loss "
        f"{info['loss_first20']:.4f} -> {info['loss_last100']:.4f}; validation (plain MG {plain:.3e}) {sel}")
    return net, info
