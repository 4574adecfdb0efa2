"""Oracle ogsSetup / gatherScatterMany pinned by brute force and by the
reference's own invariants (gs/main.cpp:136-194 degree round trip,
3rdParty/gslib/tests/gs_test.c:88-105 multiplicity asserts)."""
import numpy as np
import pytest


def brute_csr(ids):
    """ogsSetup.cpp:137-212 contract: sort (|id|, local index) with id 0
    excluded; a new group starts wherever the signed id differs from its
    predecessor (:163), so equal |id| with mixed signs splits into runs;
    groups ordered by their smallest member, members ascending."""
    nodes = sorted((abs(int(v)), i, int(v)) for i, v in enumerate(ids) if v != 0)
    groups = []
    for n, (a, i, v) in enumerate(nodes):
        if n == 0 or v != nodes[n - 1][2]:
            groups.append([])
        groups[-1].append(i)
    order = sorted(groups, key=lambda g: g[0])
    off = np.zeros(len(order) + 1, dtype=np.int32)
    off[1:] = np.cumsum([len(g) for g in order])
    return off, np.array([i for g in order for i in g], dtype=np.int32)


def test_csr_positive_ids_group_by_value(nbo):
    ids = np.array([5, 3, 0, 5, 9, 3, 3, 0, 9, 1], dtype=np.int64)
    o, off, gid, invd = nbo.ogs_setup(ids)
    try:
        assert off.tolist() == [0, 2, 5, 7, 8]
        assert gid.tolist() == [0, 3, 1, 5, 6, 4, 8, 9]
        assert invd.tolist() == [0.5, 1 / 3, 1.0, 0.5, 0.5, 1 / 3, 1 / 3, 1.0, 0.5, 1.0]
    finally:
        nbo.ogs_free(o)


def random_ids(rng, n, nlabels):
    ids = rng.integers(-nlabels, nlabels + 1, size=n).astype(np.int64)
    return ids


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_csr_matches_brute_force(nbo, seed):
    rng = np.random.default_rng(seed)
    ids = random_ids(rng, 3000, 700)
    o, off, gid, invd = nbo.ogs_setup(ids)
    try:
        boff, bids = brute_csr(ids)
        assert np.array_equal(off, boff) and np.array_equal(gid, bids)
        deg = np.diff(boff)
        for g in range(len(deg)):
            assert np.all(invd[bids[boff[g]:boff[g + 1]]] == 1.0 / deg[g])
        assert np.all(invd[ids == 0] == 1.0)
    finally:
        nbo.ogs_free(o)


def brute_gs(ids, q, k, stride, op):
    boff, bids = brute_csr(ids)
    q = q.copy()
    for f in range(k):
        for g in range(len(boff) - 1):
            m = bids[boff[g]:boff[g + 1]] + f * stride
            if len(m) < 2:
                continue
            if op == "add":
                acc = q.dtype.type(0)
                for i in m:
                    acc = acc + q[i]
            elif op == "mul":
                acc = q.dtype.type(1)
                for i in m:
                    acc = acc * q[i]
            else:
                acc = q[m[0]]
                for i in m[1:]:
                    if op == "min":
                        acc = q[i] if q[i] < acc else acc
                    else:
                        acc = q[i] if q[i] > acc else acc
            q[m] = acc
    return q


@pytest.mark.parametrize("type_,dtype", [("double", np.float64), ("float", np.float32), ("int", np.int32),
                                         ("long long int", np.int64)])
@pytest.mark.parametrize("op", ["add", "mul", "min", "max"])
def test_gs_many_matches_brute_force(nbo, type_, dtype, op):
    rng = np.random.default_rng(5)
    n, k = 400, 3
    ids = random_ids(rng, n, 120)
    stride = n + 7
    if np.issubdtype(dtype, np.integer):
        q = rng.integers(-3, 4, size=k * stride).astype(dtype)
    else:
        q = (rng.standard_normal(k * stride) * (1.0 if op != "mul" else 0.7)).astype(dtype)
    o, *_ = nbo.ogs_setup(ids)
    try:
        got = q.copy()
        nbo.gs_many(o, got, k, stride, type_, op)
        assert np.array_equal(got, brute_gs(ids, q, k, stride, op))
    finally:
        nbo.ogs_free(o)


def test_degree_round_trip_on_box(nbo):
    # gs/main.cpp:136-194: gs(1) = multiplicity, then gs(1/gs(1)) = 1 everywhere
    box = nbo.Box.make(3, 3, 4, 2, perturb=0.1, seed=1)
    m = nbo.box_mesh(box, mask_boundary=True)
    o, off, gid, invd = nbo.ogs_setup(m["ids"])
    try:
        q = np.ones(box.E * box.Np)
        nbo.gs_many(o, q, 1, q.size)
        mult = q.copy()
        masked = m["ids"] == 0
        assert set(np.unique(mult[~masked])) <= {1.0, 2.0, 4.0, 8.0}
        assert np.all(mult[masked] == 1.0)
        assert np.array_equal(1.0 / mult, invd)
        q = 1.0 / q
        nbo.gs_many(o, q, 1, q.size)
        assert np.all(q[~masked] == 1.0)
        assert np.rint(q).sum() == q.size
    finally:
        nbo.ogs_free(o)
