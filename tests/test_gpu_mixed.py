"""Mixed-precision PCG with iterative refinement (nbx_bp_pcg_mixed) against the
oracle's restatement (nbo_bp_pcg_mixed): FP32 and bf16-geometry operators
inside FP64 vectors, dots, gather-scatter and stopping test.  The operator
arithmetic and every reduction follow the same order on both sides, so the
results are compared bit for bit; the error of the mixed solution against the
FP64 solution is checked against the requested tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRECS = {"f64": 0, "f32": 1, "bf16": 2}


def H(t):
    return t.cpu().numpy()


def make(nbx, nbo, N, dims, Nf=1, bpmode=True, l1=0.0, precon=1, flexible=False):
    kw = dict(perturb=0.1, seed=40 + N)
    bp = nbx.BP(nbx.Box.make(N, *dims, **kw), Nfields=Nf, bpmode=bpmode, lambda1=l1, precon=precon,
                flexible=flexible, fused=True)
    ob = nbo.OracleBP(nbo.Box.make(N, *dims, **kw), Nf=Nf, bpmode=bpmode, lambda1=l1)
    return bp, ob


@pytest.mark.parametrize("prec", ["f32", "bf16", "f64"])
@pytest.mark.parametrize("case", [(3, (3, 3, 2), 1, True, 0.0, 1, False),
                                  (7, (2, 2, 2), 1, False, 0.5, 1, False),
                                  (4, (2, 2, 2), 3, False, 0.1, 0, True),
                                  (7, (2, 2, 2), 3, False, 0.2, 1, False),
                                  # the BP6 vector Helmholtz form of the N=9 three-field workload
                                  # (lambda0 = 1, lambda1 = 0.1, Jacobi, flexible), on 2x2x1 elements
                                  (9, (2, 2, 1), 3, False, 0.1, 1, True)],
                         ids=["poisson", "helmholtz", "bp6flex", "vector_n7", "bp6_n9"])
def test_mixed_matches_oracle(nbx, nbo, dev, prec, case):
    import torch
    N, dims, Nf, bpmode, l1, precon, flex = case
    bp, ob = make(nbx, nbo, N, dims, Nf, bpmode, l1, precon, flex)
    n = Nf * bp.fieldOffset
    inner = {"f64": 1e-4, "f32": 1e-4, "bf16": 1e-2}[prec]
    r = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(r)
    st = bp.pcg_mixed(r, x, 1e-10, 2000, prec=PRECS[prec], inner_tol=inner, max_cycles=30)
    rr, xr = ob.rhs.copy(), np.zeros(n)
    ost = ob.pcg_mixed(rr, xr, 1e-10, 2000, prec=PRECS[prec], inner_tol=inner, max_cycles=30,
                       jacobi=bool(precon), flexible=flex)
    assert st.iterations == ost.iterations and st.cycles == ost.iter
    assert st.converged == ost.converged == 1
    assert st.rdotr == ost.rdotr
    assert np.array_equal(H(x), xr)
    # error against the FP64 solution of the same system
    r64, x64 = ob.rhs.copy(), np.zeros(n)
    ob.pcg(r64, x64, 1e-13, 4000, jacobi=bool(precon), flexible=flex)
    err = np.abs(xr - x64).max() / np.abs(x64).max()
    assert err < 1e-7, err


def test_mixed_bf16_needs_more_cycles_than_f32(nbx, dev):
    import torch
    box = nbx.Box.make(5, 4, 4, 4, perturb=0.1, seed=3)
    bp = nbx.BP(box, bpmode=True, precon=nbx.PRECON_JACOBI)
    out = {}
    for name, prec, inner in (("f32", nbx.NBX_F32, 1e-5), ("bf16", nbx.NBX_BF16, 1e-2)):
        r = bp.device_array("rhs", bp.fieldOffset).clone()
        x = torch.zeros_like(r)
        st = bp.pcg_mixed(r, x, 1e-10, 5000, prec=prec, inner_tol=inner, max_cycles=40)
        assert st.converged and (st.rdotr / st.rdotr0) ** 0.5 <= 1e-10
        out[name] = st.cycles
    assert out["bf16"] > out["f32"] >= 2


def test_mixed_refuses_all_neumann(nbx, dev):
    """The singular all-Neumann problem (no Dirichlet mask, lambda1 = 0;
    nekBone/setup.cpp:210-214) has a true residual that need not reach the
    tolerance, so the refinement solver refuses it instead of running every
    cycle; nbx_bp_pcg (with BPZeroMean around it) still solves it."""
    import torch
    # nekBone's own problem: periodic box [-1,1]^3, BP5, the manufactured cosine right-hand side
    box = nbx.Box.make(5, 3, 3, 3, periodic=1, perturb=0.0, seed=5, extent=(-1.0, 1.0, -1.0, 1.0, -1.0, 1.0))
    bp = nbx.BP(box, bpmode=True, dirichlet=False, precon=nbx.PRECON_COPY, rhs_kind=1)
    assert bp.all_neumann
    r = bp.device_array("rhs", bp.fieldOffset).clone()
    x = torch.zeros_like(r)
    with pytest.raises(nbx.NbxError, match="all-Neumann"):
        bp.pcg_mixed(r, x, 1e-8, 100, prec=nbx.NBX_F32)
    st = bp.pcg(r, x, 1e-8, 1000)
    assert st.converged


@pytest.mark.parametrize("prec", ["f32", "bf16"])
def test_mixed_learned_matches_oracle_on_graded_mesh(nbx, nbo, dev, prec):
    """The stress-case solver (learned preconditioner with the FP32 or
    bf16-geometry operator inside FP64 refinement) on a small z-graded mesh
    (aspect 1:10, perturbed) against the oracle's restatement
    (nbo_bp_pcg_mixed_learned): iterations, cycles, final true residual and
    solution bit for bit, with the FP32 network inference."""
    import torch
    kw = dict(perturb=0.1, seed=44, zaspect=10.0)
    dims = (2, 2, 3)
    bp = nbx.BP(nbx.Box.make(7, *dims, **kw), bpmode=True, precon=nbx.PRECON_LEARNED)
    ob = nbo.OracleBP(nbo.Box.make(7, *dims, **kw), bpmode=True)
    net = nbx.Net(seed=3)
    W = net.params()
    W[2176:3041] = np.random.default_rng(11).uniform(-0.05, 0.05, 865).astype(np.float32)
    net.set_params(W)
    net.attach(bp)
    Wb = nbo.bf16_round(net.params())
    feat = nbo.net_features(ob.E, 512, ob.mesh["ggeo"])
    n = bp.fieldOffset
    inner = {"f32": 1e-4, "bf16": 1e-2}[prec]
    r = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(r)
    st = bp.pcg_mixed(r, x, 1e-10, 3000, prec=PRECS[prec], inner_tol=inner, max_cycles=30)
    xr = np.zeros(n)
    ost = ob.pcg_mixed_learned(ob.rhs.copy(), xr, 1e-10, 3000, Wb, feat, prec=PRECS[prec], inner_tol=inner,
                               max_cycles=30)
    print(f"learned mixed {prec}: {st.iterations} iterations, {st.cycles} cycles", flush=True)
    assert st.iterations == ost.iterations and st.cycles == ost.iter
    assert st.converged == ost.converged == 1
    assert st.rdotr == ost.rdotr
    assert np.array_equal(H(x), xr)
