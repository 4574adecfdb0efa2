"""Pins the oracle of the learned element preconditioner (oracle/nbo_net.c).

The network has no counterpart in the reference, so the pins are: an
independent float64 numpy model of the same network (forward pass and loss),
central finite differences of that model against the oracle's analytic
gradient, the zero-output-layer case (the preconditioner must be Jacobi bit
for bit) and the learned PCG with a zero network, which must reproduce the
flexible Jacobi PCG to rounding: the weighted assembly of w*invDiag*r over
m copies (w = 1/m) rounds its partial sums when m is 4 or 8."""
import numpy as np
import pytest

NP_, C_, G_, TAPS = 512, 32, 7, 27
OWR, OWG, OB1, OW2, OB2, OW3, OB3, NPAR = 0, 864, 1088, 1120, 2144, 2176, 3040, 3041


def neighbours():
    n = np.arange(NP_)
    i, j, k = n & 7, (n >> 3) & 7, n >> 6
    nb = np.full((NP_, TAPS), -1)
    for tap in range(TAPS):
        ii, jj, kk = i + tap % 3 - 1, j + (tap // 3) % 3 - 1, k + tap // 9 - 1
        ok = (ii >= 0) & (ii < 8) & (jj >= 0) & (jj < 8) & (kk >= 0) & (kk < 8)
        nb[ok, tap] = (ii + 8 * jj + 64 * kk)[ok]
    return nb


NB = neighbours()


def gather(v, nb=NB):
    """v[nb] with zero where the neighbour is outside the element."""
    out = np.where(nb >= 0, v[np.maximum(nb, 0)], 0.0)
    return out


def np_forward(W, r, g):
    """float64 model: r (512,), g (7, 512) features -> y (512,), s."""
    W = W.astype(np.float64)
    Wr = W[OWR:OWG].reshape(TAPS, C_)
    Wg = W[OWG:OB1].reshape(G_, C_)
    b1, W2, b2 = W[OB1:OW2], W[OW2:OB2].reshape(C_, C_), W[OB2:OW3]
    W3, b3 = W[OW3:OB3].reshape(TAPS, C_), W[OB3]
    s = np.abs(r).max()
    rh = r / s if s > 0 else np.zeros_like(r)
    h1 = np.maximum(b1 + g.T.astype(np.float64) @ Wg + gather(rh) @ Wr, 0.0)
    h2 = np.maximum(b2 + h1 @ W2, 0.0)
    T = h2 @ W3.T  # (512, 27)
    # y[n] = b3 + sum_tap T[nb(n, tap), tap]
    Tn = np.where(NB >= 0, T[np.maximum(NB, 0), np.arange(TAPS)[None, :]], 0.0)
    return b3 + Tn.sum(axis=1), s


def np_loss(W, samples, R, Ev, invNorm, feat, invDiag, stride):
    tot = 0.0
    for p, e in samples:
        r = R[p * stride + e * NP_: p * stride + (e + 1) * NP_].astype(np.float64)
        ev = Ev[p * stride + e * NP_: p * stride + (e + 1) * NP_].astype(np.float64)
        y, s = np_forward(W, r, feat[e * G_ * NP_:(e + 1) * G_ * NP_].reshape(G_, NP_))
        di = invDiag[e * NP_:(e + 1) * NP_]
        z = di * (r + s * y)
        tot += invNorm[p] * np.sum((z - ev) ** 2)
    return tot


def random_net(rng, scale3=0.3):
    W = np.zeros(NPAR, dtype=np.float32)
    W[OWR:OWG] = rng.uniform(-1, 1, 864) / np.sqrt(34)
    W[OWG:OB1] = rng.uniform(-1, 1, 224) / np.sqrt(34)
    W[OB1:OW2] = rng.uniform(-0.1, 0.3, 32)
    W[OW2:OB2] = rng.uniform(-1, 1, 1024) / np.sqrt(32)
    W[OB2:OW3] = rng.uniform(-0.1, 0.3, 32)
    W[OW3:OB3] = scale3 * rng.uniform(-1, 1, 864) / np.sqrt(32)
    W[OB3] = 0.05
    return W


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(11)
    E, npairs = 3, 2
    stride = E * NP_
    R = rng.standard_normal(npairs * stride).astype(np.float32)
    Ev = rng.standard_normal(npairs * stride).astype(np.float32)
    feat = rng.uniform(-1.5, 1.5, E * G_ * NP_).astype(np.float32)
    invDiag = rng.uniform(0.5, 1.5, stride)
    return dict(E=E, stride=stride, R=R, Ev=Ev, feat=feat, invDiag=invDiag, rng=rng)


def test_forward_matches_numpy_model(nbo, data):
    W = random_net(data["rng"])
    E = data["E"]
    r = data["R"][:E * NP_].astype(np.float64) * 3.0
    w = np.full(E * NP_, 0.25)
    zl = nbo.net_apply(E, W, data["feat"], r, data["invDiag"], w)
    for e in range(E):
        sl = slice(e * NP_, (e + 1) * NP_)
        y, s = np_forward(W, r[sl], data["feat"][e * G_ * NP_:(e + 1) * G_ * NP_].reshape(G_, NP_))
        zn = 0.25 * data["invDiag"][sl] * (r[sl] + s * y)
        assert np.abs(zl[sl] - zn).max() <= 2e-6 * np.abs(zn).max()


def test_zero_output_layer_is_jacobi(nbo, data):
    W = random_net(data["rng"], scale3=0.0)
    W[OB3] = 0.0
    E = data["E"]
    r = data["R"][:E * NP_].astype(np.float64)
    w = data["rng"].choice([1.0, 0.5, 0.25, 0.125], E * NP_)
    zl = nbo.net_apply(E, W, data["feat"], r, data["invDiag"], w)
    assert np.array_equal(zl, w * (data["invDiag"] * r))


def test_gradient_matches_finite_differences(nbo, data):
    rng = np.random.default_rng(5)
    W = random_net(rng)
    samples = np.array([[0, 0], [0, 2], [1, 1], [1, 2], [0, 1]], dtype=np.int32)
    invNorm = np.array([nbo.sumsq_inv_f(data["Ev"][k * data["stride"]:(k + 1) * data["stride"]]) for k in range(2)])
    g, loss = nbo.net_gradient(samples, 2, data["stride"], data["R"], data["Ev"], invNorm, W, data["feat"],
                               data["invDiag"])
    args = (samples, data["R"], data["Ev"], invNorm, data["feat"], data["invDiag"], data["stride"])
    l64 = np_loss(W, *args)
    assert abs(loss - l64) <= 1e-5 * l64
    W64 = W.astype(np.float64)
    blocks = {"Wr": (OWR, OWG), "Wg": (OWG, OB1), "b1": (OB1, OW2), "W2": (OW2, OB2), "b2": (OB2, OW3),
              "W3": (OW3, OB3), "b3": (OB3, NPAR)}
    for name, (a, b) in blocks.items():
        idx = np.arange(a, b) if b - a <= 32 else rng.choice(np.arange(a, b), 24, replace=False)
        fd = np.empty(idx.size)
        for m, q in enumerate(idx):
            h = 1e-6 * max(1.0, abs(W64[q]))  # small: the ReLU kinks bound the usable step
            Wp, Wm = W64.copy(), W64.copy()
            Wp[q] += h
            Wm[q] -= h
            fd[m] = (np_loss(Wp, *args) - np_loss(Wm, *args)) / (2 * h)
        err = np.linalg.norm(g[idx] - fd) / max(np.linalg.norm(fd), 1e-30)
        assert err < 1e-3, (name, err)  # float32 sums over 512 nodes x samples
        assert np.linalg.norm(fd) > 0, name


def np_functional(W, samples, R, G, feat, invDiag, stride):
    """sum_samples sum_n G[e, n] u_n with u = invDiag (r + s y)"""
    tot = 0.0
    for p, e in samples:
        r = R[p * stride + e * NP_: p * stride + (e + 1) * NP_].astype(np.float64)
        y, s = np_forward(W, r, feat[e * G_ * NP_:(e + 1) * G_ * NP_].reshape(G_, NP_))
        sl = slice(e * NP_, (e + 1) * NP_)
        tot += np.sum(G[sl] * invDiag[sl] * (r + s * y))
    return tot


def test_external_loss_gradient_matches_finite_differences(nbo, data):
    """G mode (energy-norm training): the gradient of sum G u over the
    parameters, u the unweighted element output."""
    rng = np.random.default_rng(15)
    W = random_net(rng)
    G = rng.standard_normal(data["E"] * NP_)
    samples = np.array([[1, 0], [0, 1], [1, 2], [0, 2]], dtype=np.int32)
    g, loss = nbo.net_gradient(samples, 3, data["stride"], data["R"], data["Ev"], np.ones(2), W, data["feat"],
                               data["invDiag"], G=G)
    assert loss == 0.0
    args = (samples, data["R"], G, data["feat"], data["invDiag"], data["stride"])
    W64 = W.astype(np.float64)
    for name, (a, b) in {"Wr": (OWR, OWG), "Wg": (OWG, OB1), "b1": (OB1, OW2), "W2": (OW2, OB2),
                         "b2": (OB2, OW3), "W3": (OW3, OB3), "b3": (OB3, NPAR)}.items():
        idx = np.arange(a, b) if b - a <= 32 else rng.choice(np.arange(a, b), 16, replace=False)
        fd = np.empty(idx.size)
        for m, q in enumerate(idx):
            h = 1e-6 * max(1.0, abs(W64[q]))
            Wp, Wm = W64.copy(), W64.copy()
            Wp[q] += h
            Wm[q] -= h
            fd[m] = (np_functional(Wp, *args) - np_functional(Wm, *args)) / (2 * h)
        err = np.linalg.norm(g[idx] - fd) / max(np.linalg.norm(fd), 1e-30)
        assert err < 1e-3, (name, err)


def test_gradient_chunking_changes_only_rounding(nbo, data):
    W = random_net(np.random.default_rng(8))
    samples = np.array([[0, 0], [1, 2], [1, 1], [0, 2]], dtype=np.int32)
    invNorm = np.array([1.0, 0.5])
    g1, l1 = nbo.net_gradient(samples, 1, data["stride"], data["R"], data["Ev"], invNorm, W, data["feat"],
                              data["invDiag"])
    g4, l4 = nbo.net_gradient(samples, 4, data["stride"], data["R"], data["Ev"], invNorm, W, data["feat"],
                              data["invDiag"])
    assert np.abs(g1 - g4).max() <= 1e-5 * np.abs(g1).max()
    assert abs(l1 - l4) <= 1e-12 * l1


def test_features_are_rms_normalised(nbo):
    box = nbo.Box.make(7, 2, 2, 1, perturb=0.1, seed=3)
    ob = nbo.OracleBP(box, bpmode=True)
    gg = ob.mesh["ggeo"].reshape(box.E, G_, NP_)
    feat = nbo.net_features(box.E, NP_, ob.mesh["ggeo"]).reshape(box.E, G_, NP_)
    rms = np.sqrt((gg ** 2).mean(axis=(0, 2)))
    assert np.allclose(feat, (gg / rms[None, :, None]).astype(np.float32), rtol=1e-6, atol=1e-7)
    assert np.allclose(np.sqrt((feat.astype(np.float64) ** 2).mean(axis=(0, 2))), 1.0, rtol=1e-6)


def test_sumsq_inv(nbo):
    a = np.random.default_rng(2).standard_normal(1000).astype(np.float32)
    assert abs(nbo.sumsq_inv_f(a) * np.sum(a.astype(np.float64) ** 2) - 1.0) < 1e-13


def test_learned_pcg_with_zero_network_is_flexible_jacobi(nbo):
    box = nbo.Box.make(7, 2, 2, 2, perturb=0.1, seed=9)
    ob = nbo.OracleBP(box, bpmode=True)
    feat = nbo.net_features(box.E, NP_, ob.mesh["ggeo"])
    W = random_net(np.random.default_rng(1), scale3=0.0)
    W[OB3] = 0.0
    n = ob.offset
    r1, x1 = ob.rhs.copy(), np.zeros(n)
    st1, h1, _ = ob.pcg_learned(r1, x1, 1e-8, 400, W, feat)
    r2, x2 = ob.rhs.copy(), np.zeros(n)
    st2, h2, _ = ob.pcg(r2, x2, 1e-8, 400, jacobi=True, flexible=True, fused=True)
    assert st1.converged and st1.iterations == st2.iterations
    assert np.abs(x1 - x2).max() <= 1e-12 * np.abs(x2).max()


def test_learned_pcg_stall_falls_back_to_jacobi(nbo):
    box = nbo.Box.make(7, 2, 2, 2, perturb=0.1, seed=9)
    ob = nbo.OracleBP(box, bpmode=True)
    feat = nbo.net_features(box.E, NP_, ob.mesh["ggeo"])
    W = random_net(np.random.default_rng(4), scale3=40.0)  # a poor preconditioner
    r, x = ob.rhs.copy(), np.zeros(ob.offset)
    st, hr, _ = ob.pcg_learned(r, x, 1e-8, 3000, W, feat, window=5, ratio=0.999)
    assert st.converged
    assert st.fallback_iter > 5
    # the stall test fired exactly at the recorded iteration
    k = st.fallback_iter
    assert not (hr[k] <= 0.999 * hr[k - 5])
    assert all(hr[i] <= 0.999 * hr[i - 5] for i in range(6, k))


def test_matrix_core_forward_tracks_fp32(nbo, data):
    """bf16-operand inference oracle: within bf16 rounding of the FP32
    forward (the correction s*y changes by about 1e-2 of itself), exactly
    Jacobi with a zero output layer."""
    W = random_net(np.random.default_rng(21))
    E = data["E"]
    r = data["R"][:E * NP_].astype(np.float64)
    w = np.full(E * NP_, 0.5)
    z32 = nbo.net_apply(E, W, data["feat"], r, data["invDiag"], w)
    zmm = nbo.net_apply_mma(E, W, data["feat"], r, data["invDiag"], w)
    jac = 0.5 * (data["invDiag"] * r)
    corr = np.abs(z32 - jac).max()
    assert corr > 0
    assert 1e-6 * corr < np.abs(zmm - z32).max() <= 5e-2 * corr
    W0 = W.copy()
    W0[OW3:NPAR] = 0.0
    assert np.array_equal(nbo.net_apply_mma(E, W0, data["feat"], r, data["invDiag"], w), jac)


def test_damped_jacobi_start_of_the_learned_multigrid_step(nbo):
    """The learned multigrid step starts from damped Jacobi
    (nekbench_amd.learned.damped_jacobi_params): with those parameters the
    network's output is y = -alpha r^, so the element output is
    (1 - alpha) w invDiag r up to the float rounding of r^ = r / s, for any
    values of the other parameters (here random), for the FP32 form; the
    matrix-core restatement rounds r^ and the hidden values to bf16, so there
    the output is that up to alpha times half a bf16 ulp of r."""
    from nekbench_amd.learned import damped_jacobi_params
    rng = np.random.default_rng(4)
    base = rng.uniform(-0.3, 0.3, NPAR).astype(np.float32)
    E = 3
    ob = nbo.OracleBP(nbo.Box.make(7, 3, 1, 1, perturb=0.1, seed=8), bpmode=True)
    feat = nbo.net_features(ob.E, ob.Np, ob.mesh["ggeo"])
    r = rng.uniform(-1, 1, E * NP_)
    invd = rng.uniform(0.5, 2.0, E * NP_)
    w = ob.invd
    for alpha in (0.5, 0.25, 0.75):
        W = nbo.bf16_round(damped_jacobi_params(base, alpha))
        exp = (1 - alpha) * w * (invd * r)
        for f, tol in ((nbo.net_apply, 1e-6), (nbo.net_apply_mma, alpha * 2.0 ** -8 / (1 - alpha))):
            z = f(E, W, feat, r, invd, w)
            assert np.abs(z - exp).max() <= tol * np.abs(exp).max(), (f.__name__, alpha)
    # the parameters are valid bf16 values where the construction sets them
    P = damped_jacobi_params(base, 0.5)
    assert np.array_equal(nbo.bf16_round(P[OW3:OB3]), P[OW3:OB3])
