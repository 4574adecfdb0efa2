"""Learned element preconditioner on the GPU against the oracle
(oracle/nbo_net.c): forward pass, loss and parameter gradient bit for bit,
the learned flexible PCG iterate for iterate (including the fall-back to
Jacobi after a stall), checkpoint round trip, and training on pairs from the
FP64 PCG lowering the loss; the trained network still solves to the FP64
solution."""
import numpy as np
import pytest

from test_oracle_net import random_net

pytestmark = pytest.mark.gpu


def H(t):
    return t.cpu().numpy()


def make(nbx, nbo, dims=(3, 3, 2), seed=21, precon=None):
    kw = dict(perturb=0.1, seed=seed)
    bp = nbx.BP(nbx.Box.make(7, *dims, **kw), bpmode=True,
                precon=nbx.PRECON_LEARNED if precon is None else precon)
    ob = nbo.OracleBP(nbo.Box.make(7, *dims, **kw), bpmode=True)
    return bp, ob


def net_with(nbx, nbo, bp, W):
    net = nbx.Net(seed=3)
    net.set_params(W)
    net.attach(bp)
    bp.update_jacobi()
    return net


def test_apply_matches_oracle(nbx, nbo, dev):
    import torch
    bp, ob = make(nbx, nbo)
    rng = np.random.default_rng(1)
    W = nbo.bf16_round(random_net(rng))
    net = net_with(nbx, nbo, bp, W)
    n = bp.fieldOffset
    invDiag = ob.invdiag()
    assert np.array_equal(H(bp.device_array("invDiagA", n)), invDiag)
    feat = nbo.net_features(ob.E, 512, ob.mesh["ggeo"])
    r = torch.tensor(rng.standard_normal(n) * 1e-3, device=dev)
    zl = torch.zeros_like(r)
    for weighted in (True, False):
        net.apply(bp, r, zl, weighted)
        exp = nbo.net_apply(ob.E, W, feat, H(r), invDiag, ob.invd if weighted else None)
        assert np.array_equal(H(zl), exp), weighted


def test_params_are_rounded_to_bf16_for_the_forward(nbx, nbo, dev):
    import torch
    bp, ob = make(nbx, nbo, dims=(2, 2, 1))
    W = random_net(np.random.default_rng(2))  # not bf16-representable
    net = net_with(nbx, nbo, bp, W)
    assert np.array_equal(net.params(), W)  # FP32 master copy kept
    r = torch.tensor(np.random.default_rng(3).standard_normal(bp.fieldOffset), device=dev)
    zl = torch.zeros_like(r)
    net.apply(bp, r, zl, True)
    feat = nbo.net_features(ob.E, 512, ob.mesh["ggeo"])
    exp = nbo.net_apply(ob.E, nbo.bf16_round(W), feat, H(r), ob.invdiag(), ob.invd)
    assert np.array_equal(H(zl), exp)


@pytest.mark.parametrize("chunk", [16, 5, 1])
def test_gradient_matches_oracle(nbx, nbo, dev, chunk):
    import torch
    bp, ob = make(nbx, nbo, dims=(2, 2, 2))
    rng = np.random.default_rng(7)
    W = nbo.bf16_round(random_net(rng))
    net = net_with(nbx, nbo, bp, W)
    n, npairs = bp.fieldOffset, 3
    R = (rng.standard_normal((npairs, n)) * 1e-2).astype(np.float32)
    Ev = (rng.standard_normal((npairs, n)) * 1e-1).astype(np.float32)
    net.set_pairs(bp, torch.tensor(R, device=dev), torch.tensor(Ev, device=dev))
    invNorm = np.array([nbo.sumsq_inv_f(Ev[k]) for k in range(npairs)])
    samples = np.stack([rng.integers(0, npairs, 37), rng.integers(0, ob.E, 37)], axis=1).astype(np.int32)
    g, loss = net.gradient(bp, samples, chunk)
    feat = nbo.net_features(ob.E, 512, ob.mesh["ggeo"])
    og, oloss = nbo.net_gradient(samples, chunk, n, R.ravel(), Ev.ravel(), invNorm, W, feat, ob.invdiag())
    assert loss == oloss
    assert np.array_equal(g, og)


def test_external_loss_gradient_matches_oracle(nbx, nbo, dev):
    import torch
    bp, ob = make(nbx, nbo, dims=(2, 2, 2))
    rng = np.random.default_rng(17)
    W = nbo.bf16_round(random_net(rng))
    net = net_with(nbx, nbo, bp, W)
    n = bp.fieldOffset
    R = (rng.standard_normal((2, n)) * 1e-2).astype(np.float32)
    Ev = np.zeros((2, n), dtype=np.float32)
    net.set_pairs(bp, torch.tensor(R, device=dev), torch.tensor(Ev + 1, device=dev))
    G = rng.standard_normal(n)
    samples = np.stack([np.ones(ob.E, np.int32), np.arange(ob.E, dtype=np.int32)], axis=1)
    g, loss = net.gradient(bp, samples, 16, G=torch.tensor(G, device=dev))
    feat = nbo.net_features(ob.E, 512, ob.mesh["ggeo"])
    og, _ = nbo.net_gradient(samples, 16, n, R.ravel(), Ev.ravel() + 1, np.ones(2), W, feat, ob.invdiag(), G=G)
    assert loss == 0.0
    assert np.array_equal(g, og)


def test_energy_training_lowers_energy_loss(nbx, dev):
    box = nbx.Box.make(7, 4, 4, 4, perturb=0.1, seed=5)
    bp = nbx.BP(box, bpmode=True, precon=nbx.PRECON_LEARNED)
    net = nbx.Net(seed=1)
    net.attach(bp)
    net.generate_pairs(bp, nseeds=4, seed0=1000, snapshots=[0, 3, 10, 30], tol_exact=1e-10)
    hist = net.train_energy(bp, steps=200, lr=2e-3, seed=3, mode=0)
    print("energy loss", hist[:3], hist[-3:])
    # at zero output layer the loss is that of one Jacobi step, below 1
    assert 0 < hist[0] < 1.0
    assert hist[-20:].mean() < 0.9 * hist[:3].mean()


def test_trained_network_cuts_iterations_on_unseen_rhs(nbx, dev):
    """What the trained network gains over the method without it (Jacobi
    PCG): pairs from seeded FP64 Jacobi solves (seeds 1000..1019), the
    scale-invariant energy loss, parameters selected on the solve of seed
    8999; the learned PCG then needs fewer iterations than Jacobi PCG on the
    right-hand sides of seeds 9000 and 9001, and reaches the same solution."""
    import torch
    box = nbx.Box.make(7, 8, 8, 8, perturb=0.1, seed=4)
    bp = nbx.BP(box, bpmode=True, precon=nbx.PRECON_LEARNED)
    bj = nbx.BP(box, bpmode=True, precon=nbx.PRECON_JACOBI, flexible=True, fused=True)
    net = nbx.Net(seed=1)
    net.attach(bp)
    net.generate_pairs(bp, nseeds=20, seed0=1000, snapshots=[0, 2, 6, 15, 35, 70, 120], tol_exact=1e-11)
    n = bp.fieldOffset
    b, r, x = (torch.empty(n, dtype=torch.float64, device=dev) for _ in range(3))
    bp.set_fallback(10 ** 6, 0.9)

    def iters(h, seed):
        h.make_rhs(seed, b)
        r.copy_(b)
        x.zero_()
        st = h.pcg(r, x, 1e-8, 4000)
        assert st.converged
        return st.iterations

    best = (iters(bp, 8999), net.params())  # zero output layer: Jacobi
    for c in range(8):
        net.train_energy(bp, steps=100, lr=1e-3, seed=7 + c, mode=1)
        its = iters(bp, 8999)
        if its < best[0]:
            best = (its, net.params())
    net.set_params(best[1])
    for seed in (9000, 9001):
        lj = iters(bj, seed)
        xj = x.clone()
        ll = iters(bp, seed)
        print("seed", seed, "jacobi", lj, "learned", ll)
        assert ll < lj
        assert (x - xj).abs().max().item() <= 1e-6 * xj.abs().max().item()


@pytest.mark.parametrize("scale3", [0.05, 0.5])
def test_learned_pcg_matches_oracle(nbx, nbo, dev, scale3):
    import torch
    bp, ob = make(nbx, nbo)
    W = nbo.bf16_round(random_net(np.random.default_rng(4), scale3=scale3))
    net = net_with(nbx, nbo, bp, W)
    n = bp.fieldOffset
    r = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(r)
    st = bp.pcg(r, x, 1e-8, 600)
    feat = nbo.net_features(ob.E, 512, ob.mesh["ggeo"])
    xr = np.zeros(n)
    ost, _, _ = ob.pcg_learned(ob.rhs.copy(), xr, 1e-8, 600, W, feat)
    assert st.iterations == ost.iterations and st.converged == ost.converged == 1
    assert st.fallback_iteration == ost.fallback_iter
    assert st.rdotr == ost.rdotr
    assert np.array_equal(H(x), xr)


def test_learned_pcg_stall_fallback_matches_oracle(nbx, nbo, dev):
    import torch
    bp, ob = make(nbx, nbo)
    W = nbo.bf16_round(random_net(np.random.default_rng(4), scale3=40.0))
    net = net_with(nbx, nbo, bp, W)
    bp.set_fallback(5, 0.999)
    n = bp.fieldOffset
    r = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(r)
    st = bp.pcg(r, x, 1e-8, 3000)
    feat = nbo.net_features(ob.E, 512, ob.mesh["ggeo"])
    xr = np.zeros(n)
    ost, _, _ = ob.pcg_learned(ob.rhs.copy(), xr, 1e-8, 3000, W, feat, window=5, ratio=0.999)
    assert ost.fallback_iter > 5 and st.fallback_iteration == ost.fallback_iter
    assert st.iterations == ost.iterations and st.converged == 1
    assert np.array_equal(H(x), xr)


def test_zero_network_is_flexible_jacobi(nbx, nbo, dev):
    import torch
    bp, _ = make(nbx, nbo)
    net = nbx.Net(seed=5)  # output layer starts at zero
    net.attach(bp)
    n = bp.fieldOffset
    r = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(r)
    st = bp.pcg(r, x, 1e-8, 600)
    bj2 = nbx.BP(nbx.Box.make(7, 3, 3, 2, perturb=0.1, seed=21), bpmode=True, precon=nbx.PRECON_JACOBI,
                 flexible=True, fused=True)
    r2 = bj2.device_array("rhs", n).clone()
    x2 = torch.zeros_like(r2)
    st2 = bj2.pcg(r2, x2, 1e-8, 600)
    assert st.iterations == st2.iterations
    assert (x - x2).abs().max().item() <= 1e-12 * x2.abs().max().item()


def test_checkpoint_round_trip(nbx, dev, tmp_path):
    net = nbx.Net(seed=9)
    W = random_net(np.random.default_rng(6))
    net.set_params(W)
    path = tmp_path / "net.nbx"
    net.save(path)
    back = nbx.Net.load(path)
    assert np.array_equal(back.params(), W)
    raw = bytearray(path.read_bytes())
    raw[100] ^= 0x10
    bad = tmp_path / "bad.nbx"
    bad.write_bytes(bytes(raw))
    with pytest.raises(nbx.NbxError):
        nbx.Net.load(bad)
    with pytest.raises(nbx.NbxError):
        nbx.Net.load(tmp_path / "missing.nbx")


def test_learned_requires_n7_single_field(nbx, dev):
    net = nbx.Net(seed=1)
    bp5 = nbx.BP(nbx.Box.make(5, 2, 2, 2), bpmode=True, precon=nbx.PRECON_LEARNED)
    with pytest.raises(nbx.NbxError):
        net.attach(bp5)
    bp3 = nbx.BP(nbx.Box.make(7, 2, 2, 2), Nfields=3, precon=nbx.PRECON_LEARNED)
    with pytest.raises(nbx.NbxError):
        net.attach(bp3)


def test_learned_without_network_fails(nbx, dev):
    import torch
    bp = nbx.BP(nbx.Box.make(7, 2, 2, 2), bpmode=True, precon=nbx.PRECON_LEARNED)
    r = bp.device_array("rhs", bp.fieldOffset).clone()
    with pytest.raises(nbx.NbxError):
        bp.pcg(r, torch.zeros_like(r), 1e-8, 10)
    # a network freed while attached detaches itself: the solve reports an error
    net = nbx.Net(seed=2)
    net.attach(bp)
    bp._net = None
    net.free()
    with pytest.raises(nbx.NbxError, match="network"):
        bp.pcg(r, torch.zeros_like(r), 1e-8, 10)
    zl = torch.zeros_like(r)
    with pytest.raises(nbx.NbxError):
        nbx.Net(seed=3).apply(bp, r, zl)


def test_training_lowers_loss(nbx, dev):
    import torch
    box = nbx.Box.make(7, 4, 4, 4, perturb=0.1, seed=5)
    bp = nbx.BP(box, bpmode=True, precon=nbx.PRECON_LEARNED)
    net = nbx.Net(seed=1)
    net.attach(bp)
    k = net.generate_pairs(bp, nseeds=6, seed0=1000, snapshots=[0, 5, 15, 40], tol_exact=1e-10)
    assert k == 24
    hist = net.train(bp, steps=300, batch=1024, lr=2e-3, seed=3)
    print("loss", hist[:3], hist[-3:])
    assert hist[-20:].mean() < 0.8 * hist[:3].mean()
    n = bp.fieldOffset
    r = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(r)
    st = bp.pcg(r, x, 1e-8, 2000)
    bj = nbx.BP(box, bpmode=True, precon=nbx.PRECON_JACOBI, fused=True)
    rj = bj.device_array("rhs", n).clone()
    xj = torch.zeros_like(rj)
    sj = bj.pcg(rj, xj, 1e-8, 2000)
    print("iterations learned", st.iterations, "jacobi", sj.iterations, "fallback", st.fallback_iteration)
    assert st.converged and sj.converged
    assert (x - xj).abs().max().item() <= 1e-6 * xj.abs().max().item()


def test_matrix_core_inference_matches_oracle(nbx, nbo, dev):
    """Inference on the matrix cores: the oracle's bf16 restatement to FP32
    rounding (the matrix cores' summation order is not specified); the
    learned PCG with it converges to the FP64 solution."""
    import torch
    bp, ob = make(nbx, nbo)
    W = nbo.bf16_round(random_net(np.random.default_rng(31), scale3=0.3))
    net = net_with(nbx, nbo, bp, W)
    net.set_inference(True)
    n = bp.fieldOffset
    r = torch.tensor(np.random.default_rng(2).standard_normal(n) * 1e-3, device=dev)
    zl = torch.zeros_like(r)
    net.apply(bp, r, zl, True)
    feat = nbo.net_features(ob.E, 512, ob.mesh["ggeo"])
    exp = nbo.net_apply_mma(ob.E, W, feat, H(r), ob.invdiag(), ob.invd)
    jac = ob.invd * ob.invdiag() * H(r)
    # a 1-ulp FP32 difference can flip a bf16 rounding of an activation: 5e-3 of the correction
    assert np.abs(H(zl) - exp).max() <= 5e-3 * np.abs(exp - jac).max()
    rr = bp.device_array("rhs", n).clone()
    x = torch.zeros_like(rr)
    st = bp.pcg(rr, x, 1e-10, 1000)
    x64 = np.zeros(n)
    ob.pcg(ob.rhs.copy(), x64, 1e-12, 3000)
    assert st.converged
    assert np.abs(H(x) - x64).max() <= 1e-8 * np.abs(x64).max()


def test_matrix_core_inference_loops_over_elements(nbx, nbo, dev):
    """More elements than the device holds workgroups at once (12 x 12 x 8 =
    1152): each workgroup computes several elements in turn, loading the next
    one's inputs while it computes the current one.  Every element matches the
    oracle's bf16 restatement as on the small mesh, the unweighted form too,
    and a repeated apply gives the same bits."""
    import torch
    bp, ob = make(nbx, nbo, dims=(12, 12, 8), seed=5)
    W = nbo.bf16_round(random_net(np.random.default_rng(37), scale3=0.3))
    net = net_with(nbx, nbo, bp, W)
    net.set_inference(True)
    n = bp.fieldOffset
    r = torch.tensor(np.random.default_rng(4).standard_normal(n) * 1e-2, device=dev)
    feat = nbo.net_features(ob.E, 512, ob.mesh["ggeo"])
    for weighted in (True, False):
        zl = torch.zeros_like(r)
        net.apply(bp, r, zl, weighted)
        z1 = H(zl)
        net.apply(bp, r, zl, weighted)
        assert np.array_equal(H(zl), z1)
        wv = ob.invd if weighted else None
        exp = nbo.net_apply_mma(ob.E, W, feat, H(r), ob.invdiag(), wv)
        jac = (ob.invd if weighted else 1.0) * ob.invdiag() * H(r)
        err = np.abs(z1 - exp).reshape(ob.E, 512).max(axis=1)
        corr = np.abs(exp - jac).max()
        print(f"weighted={weighted}: worst element {err.argmax()} error {err.max():.3e} (correction {corr:.3e})")
        assert err.max() <= 5e-3 * corr
