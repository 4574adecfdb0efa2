"""Oracle mesh, ggeo and global ids pinned by closed forms and invariants."""
import numpy as np
import pytest


def test_affine_ggeo_closed_form(nbo):
    N, NX, NY, NZ = 3, 2, 3, 4
    box = nbo.Box.make(N, NX, NY, NZ)
    m = nbo.box_mesh(box)
    Np = box.Np
    g = m["ggeo"].reshape(box.E, 7, Np)
    hx, hy, hz = 1 / NX, 1 / NY, 1 / NZ
    w = m["w"]
    W = np.einsum("k,j,i->kji", w, w, w).ravel()
    JW = hx * hy * hz / 8 * W
    assert np.allclose(g[:, 0], JW, rtol=1e-13, atol=0)
    assert np.allclose(g[:, 1], JW * (2 / hx) ** 2, rtol=1e-13)
    assert np.allclose(g[:, 4], JW * (2 / hy) ** 2, rtol=1e-13)
    assert np.allclose(g[:, 6], JW * (2 / hz) ** 2, rtol=1e-13)
    for c in (2, 3, 5):
        assert np.abs(g[:, c]).max() < 1e-12 * np.abs(g[:, 1]).max()


@pytest.mark.parametrize("perturb,zaspect", [(0.1, 0.0), (0.1, 10.0), (0.0, 4.0)])
def test_volume_and_positivity(nbo, perturb, zaspect):
    box = nbo.Box.make(4, 4, 3, 5, perturb=perturb, zaspect=zaspect, seed=3)
    m = nbo.box_mesh(box)
    g = m["ggeo"].reshape(box.E, 7, box.Np)
    # trilinear maps keep the box volume; GLL of order 2N-1 integrates J exactly
    assert abs(g[:, 0].sum() - 1.0) < 1e-13
    assert m["minJ"] > 0
    # G is symmetric positive definite at every node
    G = np.stack([np.stack([g[:, 1], g[:, 2], g[:, 3]], -1),
                  np.stack([g[:, 2], g[:, 4], g[:, 5]], -1),
                  np.stack([g[:, 3], g[:, 5], g[:, 6]], -1)], -2)
    assert np.linalg.eigvalsh(G).min() > 0
    # nodes on the box faces stay on them (boundary vertices are not perturbed)
    for c, lo, hi in (("x", 0, 1), ("y", 0, 1), ("z", 0, 1)):
        v = m[c]
        assert abs(v.min() - lo) < 1e-15 and abs(v.max() - hi) < 1e-15


def test_zgrading_aspect(nbo):
    box = nbo.Box.make(2, 8, 8, 6, zaspect=10.0)
    m = nbo.box_mesh(box)
    zc = np.unique(m["EZ"])
    dz = np.diff(zc)
    assert abs(dz.min() - (1 / 8) / 10) < 1e-12
    assert np.all(np.diff(dz) > 0)


def multiplicity_counts(N, NX, NY, NZ, periodic):
    """Local nodes by multiplicity, counted on the structured grid directly."""
    Nq = N + 1
    cnt = {}
    GX, GY, GZ = (NX * N, NY * N, NZ * N) if periodic else (NX * N + 1, NY * N + 1, NZ * N + 1)
    mult = np.zeros((GZ, GY, GX), dtype=np.int64)
    for ez in range(NZ):
        for ey in range(NY):
            for ex in range(NX):
                I = (ex * N + np.arange(Nq)) % GX if periodic else ex * N + np.arange(Nq)
                J = (ey * N + np.arange(Nq)) % GY if periodic else ey * N + np.arange(Nq)
                K = (ez * N + np.arange(Nq)) % GZ if periodic else ez * N + np.arange(Nq)
                np.add.at(mult, np.ix_(K, J, I), 1)
    for v in np.unique(mult):
        cnt[int(v)] = int((mult == v).sum() * v)
    return cnt


@pytest.mark.parametrize("periodic", [1, 0])
def test_global_id_multiplicity(nbo, periodic):
    N, NX, NY, NZ = 3, 4, 3, 5
    box = nbo.Box.make(N, NX, NY, NZ, periodic=periodic)
    m = nbo.box_mesh(box)
    ids = m["ids"]
    _, inv, counts = np.unique(ids, return_inverse=True, return_counts=True)
    mult = counts[inv]
    got = {int(v): int((mult == v).sum()) for v in np.unique(mult)}
    assert got == multiplicity_counts(N, NX, NY, NZ, periodic)
    # labels: vertices carry EToV+1 (<= Nnodes); every other group 1 + Nnodes + min local id
    Nnodes = NX * NY * NZ if periodic else (NX + 1) * (NY + 1) * (NZ + 1)
    Np = box.Np
    corner = [0, N, N + N * (N + 1), N * (N + 1)]
    corner += [c + N * (N + 1) ** 2 for c in corner]
    for e in range(box.E):
        for v, n in enumerate(corner):
            assert ids[e * Np + n] == m["EToV"][e * 8 + v] + 1
    nonvert = ids > Nnodes
    first = {}
    for i in np.nonzero(nonvert)[0]:
        first.setdefault(ids[i], i)
    assert all(lab == 1 + i + Nnodes for lab, i in first.items())


def test_periodic_16cube_histogram(nbo):
    # 16^3 elements, N=7, periodic (the reference's nekBone box): local nodes
    # by multiplicity are 216, 6*36, 12*6 and 8 per element
    ids = np.empty(4096 * 512, dtype=np.int64)
    box = nbo.Box.make(7, 16, 16, 16, periodic=1)
    nbo.lib().nbo_global_ids(nbo.C.byref(box), 0, nbo.p(ids))
    _, inv, counts = np.unique(ids, return_inverse=True, return_counts=True)
    mult = counts[inv]
    got = {int(v): int((mult == v).sum()) for v in np.unique(mult)}
    assert got == {1: 884736, 2: 884736, 4: 294912, 8: 32768}


def test_dirichlet_unknowns(nbo):
    N, NX, NY, NZ = 3, 4, 3, 2
    box = nbo.Box.make(N, NX, NY, NZ)
    m = nbo.box_mesh(box, mask_boundary=True)
    ids = m["ids"]
    assert len(np.unique(ids[ids != 0])) == (NX * N - 1) * (NY * N - 1) * (NZ * N - 1)
    o, off, gid, invd = nbo.ogs_setup(ids)
    try:
        # sum of invDegree over unmasked nodes = unique unknowns (exact dyadics)
        assert invd[ids != 0].sum() == (NX * N - 1) * (NY * N - 1) * (NZ * N - 1)
        assert np.all(invd[ids == 0] == 1.0)
    finally:
        nbo.ogs_free(o)
