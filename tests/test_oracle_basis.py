"""Oracle GLL basis and PRNG pinned by closed forms and published values."""
import numpy as np
import pytest
from numpy.polynomial import legendre as L


@pytest.mark.parametrize("N", [1, 2, 3, 5, 7, 9, 12, 15])
def test_gll_nodes_weights(nbo, N):
    z, w, D = nbo.gll(N)
    # nodes: -1, roots of P_N', +1 (numpy's Legendre root finder as independent check)
    dP = L.legder(np.eye(N + 1)[N])
    ref = np.concatenate([[-1.0], np.sort(L.legroots(dP)) if N > 1 else [], [1.0]])
    assert np.allclose(z, ref, atol=1e-14, rtol=0)
    assert np.all(np.diff(z) > 0)
    assert abs(w.sum() - 2.0) < 1e-14
    # exact for degree <= 2N-1
    for d in range(2 * N):
        exact = 0.0 if d % 2 else 2.0 / (d + 1)
        assert abs(np.dot(w, z ** d) - exact) < 1e-13
    # and not for degree 2N (GLL is not exact there)
    assert abs(np.dot(w, z ** (2 * N)) - 2.0 / (2 * N + 1)) > 1e-12


@pytest.mark.parametrize("N", [1, 2, 3, 7, 9, 15])
def test_dmatrix_exact_derivatives(nbo, N):
    z, w, D = nbo.gll(N)
    Nq = N + 1
    D = D.reshape(Nq, Nq)
    assert np.allclose(D.sum(axis=1), 0.0, atol=1e-12 * N * N)
    for d in range(N + 1):
        f = z ** d
        df = d * z ** (d - 1) if d else np.zeros_like(z)
        assert np.allclose(D @ f, df, atol=1e-11 * max(1, N) ** 2)
    assert D[0, 0] == -N * (N + 1) / 4 and D[N, N] == N * (N + 1) / 4
    # a transposed D (the classic mistake) fails the derivative test
    if N > 1:
        assert not np.allclose(D.T @ z, np.ones_like(z), atol=1e-6)


def test_splitmix64_known_answers(nbo):
    # splitmix64 (Vigna) from state 0: the first four outputs
    expect = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    golden = 0x9E3779B97F4A7C15
    got = [nbo.lib().nbo_splitmix64((golden * i) & (2 ** 64 - 1)) for i in range(4)]
    assert got == expect


def test_urand_range_and_moments(nbo):
    u = nbo.urand(7, 2, 200000)
    assert u.min() >= -1.0 and u.max() < 1.0
    assert abs(u.mean()) < 0.01 and abs(u.var() - 1.0 / 3.0) < 0.01
    # streams and seeds are independent
    assert not np.array_equal(u[:16], nbo.urand(7, 3, 16))
    assert not np.array_equal(u[:16], nbo.urand(8, 2, 16))
