"""Learned multigrid: the element network as the last step of the fine level
of every V-cycle (nbx_bp_mg_set_learned), against the oracle's restatement
(nbo_mg_vcycle with a network, OracleMG.set_net).  With the FP32 network
inference every kernel follows the oracle's order, so one V-cycle and the
multigrid-PCG iterates agree bit for bit for the FP64, FP32 and bf16-geometry
cycles, on a uniform and on a z-graded mesh.  The training pairs are checked
against their definition (A e' = r', and the first pair of a seed is the
residual one V-cycle leaves on b)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRECS = {"f64": 0, "f32": 1, "bf16": 2}


def H(t):
    return t.cpu().numpy()


def make(nbx, nbo, dims, prec, zaspect=0.0, degree=2, coarse=10, seed=61):
    kw = dict(perturb=0.1, seed=seed, zaspect=zaspect)
    bp = nbx.BP(nbx.Box.make(7, *dims, **kw), bpmode=True, precon=nbx.PRECON_MG)
    bp.mg_setup([7, 3, 1], degree, coarse, prec=prec)
    mg = nbo.OracleMG(nbo.Box.make(7, *dims, **kw), [7, 3, 1], degree=degree, coarse_iters=coarse, prec=prec)
    return bp, mg


def random_net(nbx, nbo, seed):
    """A network with a nonzero output layer (the untrained one is Jacobi)."""
    net = nbx.Net(seed=seed)
    W = net.params()
    W[2176:3041] = np.random.default_rng(seed).uniform(-0.05, 0.05, 865).astype(np.float32)
    net.set_params(W)
    return net, nbo.bf16_round(net.params())


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("zaspect", [0.0, 8.0], ids=["uniform", "graded"])
def test_learned_vcycle_and_pcg_match_oracle(nbx, nbo, dev, prec, zaspect):
    import torch
    bp, mg = make(nbx, nbo, (3, 2, 2), PRECS[prec], zaspect=zaspect)
    net, Wb = random_net(nbx, nbo, 5)
    net.attach(bp)
    bp.mg_set_learned(True)
    assert bp.mg_learned()
    mg.set_net(Wb)
    n = bp.fieldOffset
    r = mg.fine.rhs.copy()
    z = torch.zeros(n, dtype=torch.float64, device=dev)
    bp.mg_apply(torch.tensor(r, device=dev), z)
    zo = mg.vcycle(r)
    assert np.array_equal(H(z), zo)
    # the learned step changes the cycle
    mg.set_net(None)
    assert not np.array_equal(zo, mg.vcycle(r))
    mg.set_net(Wb)
    rr = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(rr)
    st = bp.pcg(rr, x, 1e-10, 200)
    xo = np.zeros(n)
    ost, _, _ = mg.pcg(mg.fine.rhs.copy(), xo, 1e-10, 200)
    assert st.iterations == ost.iterations and st.converged == ost.converged == 1
    assert st.rdotr == ost.rdotr
    assert np.array_equal(H(x), xo)


def test_learned_step_off_is_plain_multigrid(nbx, nbo, dev):
    import torch
    bp, mg = make(nbx, nbo, (2, 2, 2), 1)
    net, _ = random_net(nbx, nbo, 9)
    net.attach(bp)
    n = bp.fieldOffset
    r = torch.tensor(mg.fine.rhs, device=dev)
    z0, z1, z2 = (torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3))
    bp.mg_apply(r, z0)
    bp.mg_set_learned(True)
    bp.mg_apply(r, z1)
    bp.mg_set_learned(False)
    bp.mg_apply(r, z2)
    assert np.array_equal(H(z0), H(z2)) and not np.array_equal(H(z0), H(z1))
    assert np.array_equal(H(z0), mg.vcycle(mg.fine.rhs.copy()))


def test_learned_step_needs_a_network(nbx, dev):
    bp = nbx.BP(nbx.Box.make(7, 2, 2, 1, perturb=0.1, seed=2), bpmode=True, precon=nbx.PRECON_MG)
    bp.mg_setup([7, 3, 1], 2, 5)
    with pytest.raises(nbx.NbxError, match="no network attached"):
        bp.mg_set_learned(True)


def test_mg_pairs_are_what_a_vcycle_leaves(nbx, nbo, dev):
    """Pairs of nbx_net_generate_mg_pairs: A e' = r' (to the float rounding of
    the stored pair), and pair 0 of a seed is r' = b - A M b, e' = x* - M b."""
    import torch
    bp, mg = make(nbx, nbo, (3, 2, 2), 0)
    net, _ = random_net(nbx, nbo, 3)
    net.attach(bp)
    maxk = 3
    k = net.generate_mg_pairs(bp, nseeds=2, seed0=700, kcount=maxk, tol_exact=1e-12)
    assert k == 2 * maxk
    R, E = net.pairs(bp)
    ob = mg.fine
    for p in range(k):
        Ae = ob.apply(E[p].astype(np.float64))
        assert np.abs(Ae - R[p]).max() <= 2e-5 * np.abs(R[p]).max(), p
    # pair 0 of seed 700 from the oracle: b, x* (solved tightly), one V-cycle on b
    b = torch.empty(bp.fieldOffset, dtype=torch.float64, device=dev)
    bp.make_rhs(700, b)
    bh = H(b)
    z = mg.vcycle(bh.copy())
    rp = bh - ob.apply(z)
    assert np.abs(R[0] - rp).max() <= 1e-6 * np.abs(rp).max()
    # residuals shrink along the iterates of one seed
    nr = [float(np.abs(R[j]).max()) for j in range(maxk)]
    assert nr[maxk - 1] < nr[0]


def test_mg_pairs_from_later_iterates(nbx, nbo, dev):
    """kfirst skips the first iterates: the pairs of (kfirst, kcount) are
    pairs kfirst.. of (0, kfirst + kcount) for the same seeds."""
    import torch
    bp, mg = make(nbx, nbo, (2, 2, 2), 1)
    net, _ = random_net(nbx, nbo, 3)
    net.attach(bp)
    assert net.generate_mg_pairs(bp, nseeds=1, seed0=71, kcount=4, tol_exact=1e-12) == 4
    R0, E0 = net.pairs(bp)
    assert net.generate_mg_pairs(bp, nseeds=1, seed0=71, kcount=2, kfirst=2, tol_exact=1e-12) == 2
    R2, E2 = net.pairs(bp)
    assert np.array_equal(R2, R0[2:]) and np.array_equal(E2, E0[2:])


def test_learned_cycle_follows_the_inference_mode(nbx, nbo, dev):
    """The V-cycle is replayed from a captured HIP graph: switching the
    network's inference kernel (FP32 / matrix cores) must re-capture it, so
    every replay equals the kernel-by-kernel cycle of the current mode."""
    import torch
    bp, mg = make(nbx, nbo, (2, 2, 2), 2)
    net, _ = random_net(nbx, nbo, 13)
    net.attach(bp)
    bp.mg_set_learned(True)
    n = bp.fieldOffset
    r = torch.tensor(mg.fine.rhs, device=dev)
    out = {}
    for mode in (False, True, False):
        net.set_inference(mode)
        for graph in (True, False):
            bp.mg_use_graph(graph)
            z = torch.zeros(n, dtype=torch.float64, device=dev)
            bp.mg_apply(r, z)
            out.setdefault(mode, []).append(H(z))
        bp.mg_use_graph(True)
    for mode, zs in out.items():
        assert all(np.array_equal(zs[0], z) for z in zs[1:]), mode
    assert not np.array_equal(out[False][0], out[True][0])


def test_trained_learned_step_beats_the_same_multigrid(nbx, nbo, dev):
    """On a z-graded mesh (aspect 1:10) the network trained on the mesh's own
    multigrid pairs (nekbench_amd.learned defaults, shortened) reaches 1e-10
    in fewer multigrid-PCG iterations than the same cycle without it, on a
    right-hand side it was not trained on, with both inference kernels; with
    the FP32 inference the solve follows the oracle's iterates bit for bit
    for the trained weights."""
    import torch
    from nekbench_amd.learned import train_mg_network
    dims, kw = (8, 8, 4), dict(perturb=0.1, seed=2, zaspect=10.0)
    bp = nbx.BP(nbx.Box.make(7, *dims, **kw), bpmode=True, precon=nbx.PRECON_MG)
    bp.mg_setup([7, 3, 1], 3, 8, prec=PRECS["f32"])
    n = bp.fieldOffset
    b = torch.empty(n, dtype=torch.float64, device=dev)
    r, x = torch.empty_like(b), torch.empty_like(b)
    bp.make_rhs(4242, b)

    def solve():
        r.copy_(b)
        x.zero_()
        return bp.pcg(r, x, 1e-10, 300)

    plain = solve()
    net, info = train_mg_network(nbx, torch, bp, seeds=(1000, 24), steps=1000, log=print)
    its = {}
    for mfma in (False, True):
        net.set_inference(mfma)
        st = solve()
        assert st.converged
        its[mfma] = st.iterations
    print(f"plain {plain.iterations} iterations, learned fp32 {its[False]}, matrix cores {its[True]}")
    assert plain.converged and its[False] < plain.iterations and its[True] < plain.iterations
    # the trained weights in the oracle's cycle: the same iterates
    net.set_inference(False)
    mg = nbo.OracleMG(nbo.Box.make(7, *dims, **kw), [7, 3, 1], degree=3, coarse_iters=8, prec=PRECS["f32"])
    mg.set_net(nbo.bf16_round(net.params()))
    st = solve()
    xo = np.zeros(n)
    ost, _, _ = mg.pcg(H(b), xo, 1e-10, 300)
    assert st.iterations == ost.iterations and st.rdotr == ost.rdotr
    assert np.array_equal(H(x), xo)
