"""p-multigrid preconditioner on the GPU against the oracle (OracleMG):
power-iteration lambda_max per level, one V-cycle, and the MG-preconditioned
flexible PCG iterate for iterate, all bit for bit, for the FP64 cycle and the
FP32 / bf16-geometry cycles; argument checks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def H(t):
    return t.cpu().numpy()


CASES = [  # (N, dims, orders, degree, coarse_iters, bpmode, lambda1)
    (7, (3, 3, 2), [7, 3, 1], 2, 10, True, 0.0),
    (5, (2, 3, 2), [5, 2, 1], 3, 6, False, 0.7),
    (3, (3, 2, 2), [3, 1], 1, 12, True, 0.0),
    # the BP6 orders (transfers compiled for 10->5, 5->2); one element layer: every order-1 vertex lies
    # on the Dirichlet boundary, so the coarse solve has no unknowns
    (9, (2, 2, 1), [9, 4, 1], 2, 8, False, 0.3),
    (7, (1, 1, 1), [7, 3, 1], 2, 10, True, 0.0),  # one element: no shared node on any level
    (5, (2, 3, 2), [5, 2, 1], 2, 8, False, 0.5, 1),  # periodic Helmholtz: no mask, every coarse vertex shared
    (5, (2, 2, 2), [5, 2], 2, 12, True, 0.0),  # last level of order 2: the generic five-kernel coarse PCG
]


def make(nbx, nbo, case, prec=0):
    N, dims, orders, deg, ci, bpmode, l1 = case[:7]
    kw = dict(perturb=0.1, seed=50 + N, periodic=case[7] if len(case) > 7 else 0)
    bp = nbx.BP(nbx.Box.make(N, *dims, **kw), bpmode=bpmode, lambda1=l1, precon=nbx.PRECON_MG)
    bp.mg_setup(orders, deg, ci, prec=prec)
    mg = nbo.OracleMG(nbo.Box.make(N, *dims, **kw), orders, degree=deg, coarse_iters=ci, bpmode=bpmode,
                      lambda1=l1, prec=prec)
    return bp, mg


PRECS = {"f64": 0, "f32": 1, "bf16": 2}


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("case", CASES, ids=["n7", "n5helm", "n3", "n9helm", "n7one", "n5per", "n5to2"])
def test_mg_setup_and_vcycle_match_oracle(nbx, nbo, dev, case, prec):
    import torch
    bp, mg = make(nbx, nbo, case, PRECS[prec])
    orders = case[2]
    for l in range(len(orders) - 1):
        N, lam = bp.mg_info(l)
        assert N == orders[l]
        assert lam == mg.lmax[l]
    n = bp.fieldOffset
    r = mg.fine.rhs.copy()
    z = torch.zeros(n, dtype=torch.float64, device=dev)
    bp.mg_apply(torch.tensor(r, device=dev), z)
    assert np.array_equal(H(z), mg.vcycle(r))


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("case", CASES, ids=["n7", "n5helm", "n3", "n9helm", "n7one", "n5per", "n5to2"])
def test_mg_pcg_iterates_match_oracle(nbx, nbo, dev, case, prec):
    import torch
    bp, mg = make(nbx, nbo, case, PRECS[prec])
    n = bp.fieldOffset
    r = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(r)
    st = bp.pcg(r, x, 1e-10, 200)
    xr = np.zeros(n)
    ost, _, _ = mg.pcg(mg.fine.rhs.copy(), xr, 1e-10, 200)
    assert st.iterations == ost.iterations and st.converged == ost.converged == 1
    assert st.rdotr == ost.rdotr
    assert np.array_equal(H(x), xr)


def test_mg_bf16_mfma_cycle(nbx, nbo, dev):
    """V-cycle with the N=7 level on the matrix cores: close to the oracle's
    restatement (the matrix cores' summation order is not specified, so not
    bit for bit), and the MG-PCG reaches the FP64 solution in the oracle's
    iteration count (+-1)."""
    import torch
    case = CASES[0]
    N, dims, orders, deg, ci, bpmode, l1 = case
    kw = dict(perturb=0.1, seed=50 + N)
    bp = nbx.BP(nbx.Box.make(N, *dims, **kw), bpmode=bpmode, lambda1=l1, precon=nbx.PRECON_MG)
    bp.mg_setup(orders, deg, ci, prec=nbx.NBX_BF16_MFMA)
    mg = nbo.OracleMG(nbo.Box.make(N, *dims, **kw), orders, degree=deg, coarse_iters=ci, bpmode=bpmode,
                      lambda1=l1, prec=3)
    n = bp.fieldOffset
    r = mg.fine.rhs.copy()
    z = torch.zeros(n, dtype=torch.float64, device=dev)
    bp.mg_apply(torch.tensor(r, device=dev), z)
    zo = mg.vcycle(r)
    assert np.abs(H(z) - zo).max() <= 1e-3 * np.abs(zo).max()
    rr = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(rr)
    st = bp.pcg(rr, x, 1e-10, 200)
    xo = np.zeros(n)
    ost, _, _ = mg.pcg(mg.fine.rhs.copy(), xo, 1e-10, 200)
    assert st.converged and abs(st.iterations - ost.iterations) <= 1
    assert np.abs(H(x) - xo).max() <= 1e-8 * np.abs(xo).max()


def test_mg_cuts_iterations(nbx, dev):
    import torch
    box = nbx.Box.make(7, 6, 6, 6, perturb=0.1, seed=3)
    bp = nbx.BP(box, bpmode=True, precon=nbx.PRECON_MG)
    bp.mg_setup([7, 3, 1], 2, 20)
    bj = nbx.BP(box, bpmode=True, precon=nbx.PRECON_JACOBI, fused=True)
    n = bp.fieldOffset
    r = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(r)
    st = bp.pcg(r, x, 1e-8, 500)
    rj = bj.device_array("rhs", n).clone()
    xj = torch.zeros_like(rj)
    sj = bj.pcg(rj, xj, 1e-8, 2000)
    print("mg", st.iterations, st.elapsed_s, "jacobi", sj.iterations, sj.elapsed_s)
    assert st.converged and sj.converged and st.iterations * 10 < sj.iterations
    assert (x - xj).abs().max().item() <= 1e-6 * xj.abs().max().item()


def test_mg_argument_checks(nbx, dev):
    import torch
    bp = nbx.BP(nbx.Box.make(5, 2, 2, 2), bpmode=True, precon=nbx.PRECON_MG)
    r = bp.device_array("rhs", bp.fieldOffset).clone()
    with pytest.raises(nbx.NbxError):
        bp.pcg(r, torch.zeros_like(r), 1e-8, 10)  # not set up
    for orders in ([4, 2, 1], [5, 5, 1], [5, 2, 0], [5]):
        with pytest.raises(nbx.NbxError):
            bp.mg_setup(orders)
    with pytest.raises(nbx.NbxError):
        bp.mg_setup([5, 2, 1], prec=nbx.NBX_F16)
    bn = nbx.BP(nbx.Box.make(5, 3, 3, 3, periodic=1), bpmode=True, dirichlet=False, precon=nbx.PRECON_MG)
    with pytest.raises(nbx.NbxError, match="all-Neumann"):
        bn.mg_setup([5, 2, 1])
    b3 = nbx.BP(nbx.Box.make(5, 2, 2, 2), Nfields=3, precon=nbx.PRECON_MG)
    with pytest.raises(nbx.NbxError):
        b3.mg_setup([5, 2, 1], prec=nbx.NBX_BF16)  # bf16 geometry: one field


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_mg_three_fields_match_oracle(nbx, nbo, dev, prec):
    """BP6 (three fields, non-BK Helmholtz): the cycle acts field by field
    with one diagonal and one lambda_max; V-cycle and PCG iterates bit for bit."""
    import torch
    kw = dict(perturb=0.1, seed=61)
    bp = nbx.BP(nbx.Box.make(4, 3, 3, 2, **kw), Nfields=3, bpmode=False, lambda1=0.1, precon=nbx.PRECON_MG)
    bp.mg_setup([4, 2, 1], 2, 12, prec=PRECS[prec])
    mg = nbo.OracleMG(nbo.Box.make(4, 3, 3, 2, **kw), [4, 2, 1], degree=2, coarse_iters=12, bpmode=False,
                      lambda1=0.1, prec=PRECS[prec], Nf=3)
    n = 3 * bp.fieldOffset
    r = mg.fine.rhs.copy()
    z = torch.zeros(n, dtype=torch.float64, device=dev)
    bp.mg_apply(torch.tensor(r, device=dev), z)
    assert np.array_equal(H(z), mg.vcycle(r))
    rr = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(rr)
    st = bp.pcg(rr, x, 1e-10, 200)
    xo = np.zeros(n)
    ost, _, _ = mg.pcg(mg.fine.rhs.copy(), xo, 1e-10, 200)
    assert st.iterations == ost.iterations and st.converged
    assert np.array_equal(H(x), xo)


def test_mg_graph_replay_equals_direct_launches(nbx, dev):
    """The cycle replayed from its HIP graph gives the bits of the cycle
    launched kernel by kernel, over a whole solve and after a change of the
    residual vector (which re-captures the graph)."""
    import torch
    box = nbx.Box.make(7, 4, 3, 3, perturb=0.1, seed=13)
    bp = nbx.BP(box, bpmode=True, precon=nbx.PRECON_MG)
    bp.mg_setup([7, 3, 1], 2, 15, prec=nbx.NBX_F32)
    n = bp.fieldOffset
    out = {}
    for on in (True, False):
        bp.mg_use_graph(on)
        for k in range(2):  # a second residual buffer forces a new capture
            r = bp.device_array("rhs", n).clone()
            x = torch.zeros_like(r)
            st = bp.pcg(r, x, 1e-10, 300)
            out[(on, k)] = (st.iterations, x.cpu())
    for k in range(2):
        assert out[(True, k)][0] == out[(False, k)][0]
        assert torch.equal(out[(True, k)][1], out[(False, k)][1])
