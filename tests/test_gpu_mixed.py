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
                                  (7, (2, 2, 2), 3, False, 0.2, 1, False)],
                         ids=["poisson", "helmholtz", "bp6flex", "vector_n7"])
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
