"""GPU parity of the BP driver (nbx_bp_*): setup products, BPAxOperator and
every PCG iterate against the oracle's restatement of BPPCG; full-size
properties on the 32^3-element N=7 Dirichlet box."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def H(t):
    return t.cpu().numpy()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# The update pass has a one-field path for full 2048-node tiles besides its
# general loop; the one-field cases cover local node counts with only full
# tiles (case1: 6144), none (case0: 1152, case9: 512) and both (case2: 2592,
# case5: 5832), with the copy, Jacobi and flexible variants.
CASES = [
    # (N, dims, Nf, bpmode, lambda1, precon, flexible, periodic, rhs_kind)
    (3, (3, 3, 2), 1, False, 0.0, 1, False, 0, 0),     # Jacobi Poisson, Dirichlet
    (7, (3, 2, 2), 1, True, 0.0, 0, False, 0, 0),      # BP5, copy preconditioner
    (5, (2, 3, 2), 1, False, 0.7, 1, True, 0, 0),      # Helmholtz, Jacobi, flexible
    (4, (2, 2, 2), 3, True, 0.0, 0, False, 0, 0),      # BP6 (3 fields)
    (3, (3, 4, 3), 3, False, 0.3, 0, True, 0, 0),      # vector Helmholtz, flexible
    (5, (3, 3, 3), 1, False, 1.1, 1, False, 1, 1),     # periodic manufactured (reference nekBone)
    (7, (2, 2, 2), 3, False, 0.3, 1, True, 0, 0),      # vector Helmholtz N=7: one thread group per field
    (9, (2, 2, 1), 3, False, 0.2, 0, False, 0, 0),     # vector Helmholtz N=9: field groups, D from LDS
    (7, (2, 2, 2), 3, True, 0.0, 0, False, 0, 0),      # BP6 N=7: field loop
    (7, (1, 1, 1), 1, False, 0.0, 1, False, 0, 0),     # one element: no shared node, the mask alone
]


def make(nbx, nbo, case, fused=True):
    N, dims, Nf, bpmode, l1, precon, flex, periodic, rhs_kind = case
    kw = dict(perturb=0.0 if periodic else 0.1, seed=31 + N, periodic=periodic)
    if periodic:
        kw["extent"] = (-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)
    bp = nbx.BP(nbx.Box.make(N, *dims, **kw), Nfields=Nf, bpmode=bpmode, lambda1=l1, dirichlet=not periodic,
                precon=precon, flexible=flex, rhs_kind=rhs_kind, fused=fused)
    ob = nbo.OracleBP(nbo.Box.make(N, *dims, **kw), Nf=Nf, bpmode=bpmode, lambda1=l1, dirichlet=not periodic,
                      rhs_kind=rhs_kind)
    return bp, ob


@pytest.mark.parametrize("case", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_bp_setup_and_operator(nbx, nbo, dev, case):
    import torch
    bp, ob = make(nbx, nbo, case)
    n = bp.Nfields * bp.fieldOffset
    rhs = H(bp.device_array("rhs", n))
    if case[-1] == 0:
        assert np.array_equal(rhs, ob.rhs)
    else:  # host libm cos vs numpy cos
        assert rel(rhs, ob.rhs) < 1e-14
    assert np.array_equal(H(bp.device_array("invDegree", bp.Nlocal)), ob.invd)
    q = torch.from_numpy(ob.rhs.copy()).to(dev)
    Aq = torch.empty_like(q)
    pAp = bp.ax_operator(q, Aq)
    ref = ob.apply(ob.rhs.copy())
    assert np.array_equal(H(Aq), ref)
    assert pAp == ob.wdot(ob.rhs.copy(), ref)


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "refsched"])
@pytest.mark.parametrize("case", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_bp_pcg_iterates_match_oracle(nbx, nbo, dev, case, fused):
    import torch
    bp, ob = make(nbx, nbo, case, fused)
    precon, flex = case[5], case[6]
    n = bp.Nfields * bp.fieldOffset
    maxit, tol = 60, 1e-10
    r = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(r)
    st = bp.pcg(r, x, tol, maxit)
    hr, ha = bp.history(maxit + 2)
    rr, xr = ob.rhs.copy(), np.zeros(n)
    ost, ohr, oha = ob.pcg(rr, xr, tol, maxit, jacobi=bool(precon), flexible=flex, fused=fused)
    assert st.iterations == ost.iterations and st.converged == ost.converged
    k = st.iterations + 1
    assert np.array_equal(hr[1:k + 1], ohr[1:k + 1]), "rdotr history differs"
    assert np.array_equal(ha[1:k], oha[1:k]), "alpha history differs"
    assert rel(H(x), xr) <= 1e-12 and rel(H(r), rr) <= 1e-12
    assert np.array_equal(H(x), xr) and np.array_equal(H(r), rr)


def test_bp_pcg_fixed_iterations_and_restart(nbx, nbo, dev):
    import torch
    bp, ob = make(nbx, nbo, CASES[0])
    n = bp.fieldOffset
    r = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(r)
    st = bp.pcg(r, x, 0.0, 7)  # tol <= 0: FIXED ITERATION COUNT
    assert st.iterations == 7 and not st.converged
    rr, xr = ob.rhs.copy(), np.zeros(n)
    ob.pcg(rr, xr, 0.0, 7)
    assert np.array_equal(H(x), xr)
    # second solve on the same handle continues from x (r reset by the caller)
    r2 = bp.device_array("rhs", n).clone()
    st2 = bp.pcg(r2, x, 1e-9, 200)
    rr2 = ob.rhs.copy()
    ost2, _, _ = ob.pcg(rr2, xr, 1e-9, 200)
    assert st2.iterations == ost2.iterations and np.array_equal(H(x), xr)


def test_bp_pcg_zero_rhs_converges_immediately(nbx, nbo, dev):
    import torch
    bp, _ = make(nbx, nbo, CASES[0])
    r = torch.zeros(bp.fieldOffset, dtype=torch.float64, device=dev)
    x = torch.zeros_like(r)
    st = bp.pcg(r, x, 1e-8, 50)
    assert st.iterations == 0 and st.converged and float(x.abs().max()) == 0.0


def test_manufactured_solution_error(nbx, dev):
    # nekBone's correctness check (main.cpp:105-132): periodic [-1,1]^3,
    # Helmholtz lambda=1.1, exact solution cos(pi x)cos(pi y)cos(pi z)
    import torch
    errs = []
    for N in (3, 5, 7):
        box = nbx.Box.make(N, 4, 4, 4, periodic=1, extent=(-1.0, 1.0, -1.0, 1.0, -1.0, 1.0))
        bp = nbx.BP(box, lambda1=1.1, dirichlet=False, precon=nbx.PRECON_JACOBI, rhs_kind=1)
        r = bp.device_array("rhs", bp.fieldOffset).clone()
        x = torch.zeros_like(r)
        st = bp.pcg(r, x, 1e-12, 2000)
        assert st.converged
        m = nbx.box_mesh(box, nodes=True, ids=False)
        exact = np.cos(np.pi * m["x"]) * np.cos(np.pi * m["y"]) * np.cos(np.pi * m["z"])
        errs.append(float(np.abs(H(x) - exact).max()))
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-4, errs


def test_32cube_full_size_properties(nbx, dev):
    """32^3 elements, N=7, Dirichlet, perturbed: Jacobi PCG to 1e-8 or 500
    iterations; the final residual is checked by recomputing r = b - A x."""
    import torch
    box = nbx.Box.make(7, 32, 32, 32, perturb=0.1, seed=1)
    bp = nbx.BP(box, precon=nbx.PRECON_JACOBI)
    n = bp.fieldOffset
    assert bp.Nlocal == 32 ** 3 * 512
    b = bp.device_array("rhs", n).clone()
    r, x = b.clone(), torch.zeros_like(b)
    st = bp.pcg(r, x, 1e-8, 500)
    assert st.iterations <= 500 and st.rdotr <= st.rdotr0
    Ax = torch.empty_like(x)
    bp.ax_operator(x, Ax)
    true_r = b - Ax
    w = bp.device_array("invDegree", n)
    rr_true = float((w * true_r * true_r).sum())
    assert abs(rr_true - st.rdotr) <= 1e-6 * st.rdotr0 + 1e-3 * st.rdotr
    if st.converged:
        assert st.rdotr <= 1e-16 * st.rdotr0


@pytest.mark.parametrize("Nf", [1, 3])
def test_all_neumann_zero_mean_matches_oracle(nbx, nbo, dev, Nf):
    """lambda1 = 0 without a mask: BPZeroMean on r before and on x after the
    solve (nekBone/main.cpp:24-31); fixed iteration count as nekBone runs it."""
    import torch
    kw = dict(perturb=0.0, seed=5, periodic=1)
    bp = nbx.BP(nbx.Box.make(4, 3, 3, 3, **kw), Nfields=Nf, bpmode=True, dirichlet=False,
                precon=nbx.PRECON_COPY)
    ob = nbo.OracleBP(nbo.Box.make(4, 3, 3, 3, **kw), Nf=Nf, bpmode=True, dirichlet=False)
    assert bp.all_neumann and ob.all_neumann
    n = Nf * bp.fieldOffset
    q = torch.tensor(np.random.default_rng(2).uniform(-1, 1, n) + 0.25, device=dev)
    qo = H(q).copy()
    bp.zero_mean(q)
    ob.zero_mean(qo)
    torch.cuda.synchronize()
    assert np.array_equal(H(q), qo)
    r = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(r)
    st = bp.pcg(r, x, 0.0, 25)
    rr, xr = ob.rhs.copy(), np.zeros(n)
    ost, _, _ = ob.pcg(rr, xr, 0.0, 25, jacobi=False)
    assert st.iterations == ost.iterations == 25
    assert np.array_equal(H(x), xr)
    assert np.array_equal(H(r), rr)
    assert abs(xr.sum()) <= 1e-12 * np.abs(xr).sum()


def test_bp6_jacobi_matches_oracle(nbx, nbo, dev):
    """JACOBI with three fields (one diagonal shared by the fields of the
    vector operator): iterates equal the oracle's."""
    import torch
    kw = dict(perturb=0.1, seed=8)
    bp = nbx.BP(nbx.Box.make(4, 3, 2, 2, **kw), Nfields=3, lambda1=0.3, precon=nbx.PRECON_JACOBI, flexible=True,
                fused=True)
    ob = nbo.OracleBP(nbo.Box.make(4, 3, 2, 2, **kw), Nf=3, lambda1=0.3)
    n = 3 * bp.fieldOffset
    r = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(r)
    st = bp.pcg(r, x, 1e-10, 400)
    xr = np.zeros(n)
    ost, _, _ = ob.pcg(ob.rhs.copy(), xr, 1e-10, 400, jacobi=True, flexible=True)
    assert st.converged and st.iterations == ost.iterations
    assert np.array_equal(H(x), xr)
