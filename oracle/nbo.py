"""ctypes view of the CPU oracle (libnbo.so) and of the reference CPU kernels
compiled into oracle/_ref.  TEST INFRASTRUCTURE ONLY: imported by tests/,
the smoke check and bench.py's CPU-baseline leg, never by the product."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF_DIR = HERE / "_ref"
_lib = None


class Box(C.Structure):
    _fields_ = [("N", C.c_int), ("NX", C.c_int), ("NY", C.c_int), ("NZ", C.c_int),
                ("periodic", C.c_int), ("perturb", C.c_double), ("zaspect", C.c_double),
                ("seed", C.c_uint64),
                ("x0", C.c_double), ("x1", C.c_double), ("y0", C.c_double), ("y1", C.c_double),
                ("z0", C.c_double), ("z1", C.c_double)]

    @classmethod
    def make(cls, N, NX, NY, NZ, periodic=0, perturb=0.0, zaspect=0.0, seed=1,
             extent=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)):
        return cls(N, NX, NY, NZ, periodic, perturb, zaspect, seed, *extent)

    @property
    def E(self):
        return self.NX * self.NY * self.NZ

    @property
    def Np(self):
        return (self.N + 1) ** 3


class Ogs(C.Structure):
    _fields_ = [("N", C.c_int32), ("Nlocal", C.c_int32), ("NlocalGather", C.c_int32),
                ("offsets", C.POINTER(C.c_int32)), ("gatherIds", C.POINTER(C.c_int32)),
                ("invDegree", C.POINTER(C.c_double))]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        # NBO_LIB selects another build of the same sources (the AddressSanitizer
        # one of `make -C oracle asan`, run by tests/test_asan.py)
        path = Path(os.environ["NBO_LIB"]) if os.environ.get("NBO_LIB") else HERE / "libnbo.so"
        if not path.exists():
            raise RuntimeError(f"{path} not built (make -C oracle)")
        _lib = C.CDLL(str(path))
        _lib.nbo_urand.restype = C.c_double
        _lib.nbo_urand.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        _lib.nbo_splitmix64.restype = C.c_uint64
        _lib.nbo_splitmix64.argtypes = [C.c_uint64]
        _lib.nbo_geometric_factors.restype = C.c_double
        for f in ("nbo_tree_wdot", "nbo_tree_dot", "nbo_tree_sum_partials"):
            getattr(_lib, f).restype = C.c_double
        _lib.nbo_box_nelements.restype = C.c_int64
    return _lib


def p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    return C.c_void_p(a.ctypes.data)


def urand(seed, stream, n):
    out = np.empty(n)
    lib().nbo_urand_fill(C.c_uint64(seed), C.c_uint64(stream), C.c_int64(n), p(out))
    return out


def gll(N):
    z = np.empty(N + 1)
    w = np.empty(N + 1)
    lib().nbo_gll(N, p(z), p(w))
    D = np.empty((N + 1) ** 2)
    lib().nbo_dmatrix(N, p(z), p(D))
    return z, w, D


def box_mesh(box: Box, mask_boundary=False):
    """Oracle mesh: dict with EX/EY/EZ, EToV, x, y, z, ggeo, ids, D, w, z_gll."""
    E, N, Np = box.E, box.N, box.Np
    l = lib()
    EX, EY, EZ = (np.empty(E * 8) for _ in range(3))
    EToV = np.empty(E * 8, dtype=np.int64)
    l.nbo_box_vertices(C.byref(box), p(EX), p(EY), p(EZ), p(EToV))
    zg, wg, D = gll(N)
    x, y, z = (np.empty(E * Np) for _ in range(3))
    l.nbo_physical_nodes(N, C.c_int64(E), p(EX), p(EY), p(EZ), p(zg), p(x), p(y), p(z))
    ggeo = np.empty(E * 7 * Np)
    minJ = l.nbo_geometric_factors(N, C.c_int64(E), p(x), p(y), p(z), p(D), p(wg), p(ggeo))
    ids = np.empty(E * Np, dtype=np.int64)
    l.nbo_global_ids(C.byref(box), int(mask_boundary), p(ids))
    return dict(EX=EX, EY=EY, EZ=EZ, EToV=EToV, x=x, y=y, z=z, ggeo=ggeo, ids=ids, D=D, w=wg,
                z_gll=zg, minJ=minJ)


def ogs_setup(ids):
    ids = np.ascontiguousarray(ids, dtype=np.int64)
    o = Ogs()
    lib().nbo_ogs_setup(C.c_int32(ids.size), p(ids), C.byref(o))
    off = np.ctypeslib.as_array(o.offsets, shape=(o.NlocalGather + 1,)).copy()
    gid = np.ctypeslib.as_array(o.gatherIds, shape=(max(o.Nlocal, 1),)).copy()[: o.Nlocal]
    invd = np.ctypeslib.as_array(o.invDegree, shape=(max(o.N, 1),)).copy()[: o.N]
    return o, off, gid, invd


def ogs_free(o):
    lib().nbo_ogs_free(C.byref(o))


TYPES = {"double": 0, "float": 1, "int": 2, "long long int": 3}
OPS = {"add": 0, "mul": 1, "min": 2, "max": 3}


def gs_many(o, q, k, stride, type_="double", op="add"):
    rc = lib().nbo_gs_many(C.byref(o), p(q), k, C.c_int32(stride), TYPES[type_], OPS[op])
    assert rc == 0


def axhelm(N, Nf, bk, E, offset, ggeo, D, lam, q, elementList=None, out=None):
    Aq = np.zeros_like(q) if out is None else out
    lib().nbo_axhelm(N, Nf, int(bk), C.c_int32(E), C.c_int32(offset), p(elementList), p(ggeo), p(D),
                     p(lam), p(q), p(Aq))
    return Aq


def axhelm_f32(N, Nf, bk, E, offset, ggeo, D, lam, q):
    Aq = np.zeros_like(q)
    lib().nbo_axhelm_f32(N, Nf, int(bk), C.c_int32(E), C.c_int32(offset), p(ggeo), p(D), p(lam), p(q), p(Aq))
    return Aq


def axhelm_reference(Nq, E, lam, ggeo, D, q):
    out = np.zeros_like(q)
    lib().nbo_axhelm_reference(Nq, E, C.c_double(lam), p(ggeo), p(D), p(q), p(out))
    return out


def jacobi_local(N, E, offset, ggeo, D, lam, mapB=None):
    out = np.zeros(E * (N + 1) ** 3)
    lib().nbo_jacobi_local(N, C.c_int32(E), C.c_int32(offset), p(mapB), p(ggeo), p(D), p(lam), p(out))
    return out


def tree_wdot(w, x, y):
    return lib().nbo_tree_wdot(C.c_int64(x.size), p(w), p(x), p(y))


def tree_dot(x, y):
    return lib().nbo_tree_dot(C.c_int64(x.size), p(x), p(y))


# ---------------------------------------------------------------- oracle/_ref
def ref_available(N=7) -> bool:
    return (REF_DIR / f"libaxhelm_N{N}.so").exists() and (REF_DIR / "libaxref.so").exists()


_ref_cache = {}


def _ref(name):
    if name not in _ref_cache:
        _ref_cache[name] = C.CDLL(str(REF_DIR / name))
    return _ref_cache[name]


def ref_axhelm(N, kernel, E, offset, ggeo, D, lam, q, out=None):
    """Call the reference's compiled CPU kernel axhelm[_bk][_n3]_v0 (axhelm/kernel/CPU/axhelm.c)."""
    l = _ref(f"libaxhelm_N{N}.so")
    f = getattr(l, kernel)
    Aq = np.zeros_like(q) if out is None else out
    if lam is None:
        lam = np.zeros(2 * offset)
    f(C.byref(C.c_int(E)), C.byref(C.c_int(offset)), p(ggeo), p(D), p(lam), p(q), p(Aq))
    return Aq


def ref_axhelm_reference(Nq, E, lam, ggeo, D, q):
    l = _ref("libaxref.so")
    out = np.zeros_like(q)
    l.ref_axhelmReference(Nq, E, C.c_double(lam), p(ggeo), p(D), p(q), p(out))
    return out


def ref_weighted_inner_product_update(Nf, N, offset, w, a, b):
    """weightedInnerProductUpdate (nekBone/kernel/CPU/weightedInnerProduct.c:63-84)."""
    l = _ref(f"libwip_nf{Nf}.so")
    out = np.zeros(2)
    l.weightedInnerProductUpdate(C.byref(C.c_int(N)), C.byref(C.c_int(offset)), C.byref(C.c_int(1)),
                                 p(w), p(a), p(b), p(out))
    return out


class BPStruct(C.Structure):
    _fields_ = [("N", C.c_int), ("Nf", C.c_int), ("bk", C.c_int), ("E", C.c_int32), ("offset", C.c_int32),
                ("ggeo", C.c_void_p), ("D", C.c_void_p), ("lam", C.c_void_p), ("ogs", C.c_void_p),
                ("w", C.c_void_p), ("Nmasked", C.c_int32), ("maskIds", C.c_void_p)]


class PCGStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("breakdown", C.c_int), ("iter", C.c_int),
                ("rdotr0", C.c_double), ("rdotr", C.c_double), ("TOL", C.c_double), ("fallback_iter", C.c_int)]


class OracleBP:
    """The oracle's view of nekBone's BP problem: mesh, ogs, lambda, rhs --
    built exactly as the library's nbx_bp_setup builds them."""

    def __init__(self, box: Box, Nf=1, bpmode=False, lambda1=0.0, dirichlet=True, rhs_kind=0):
        self.box = box
        self.N, self.Nf = box.N, Nf
        m = box_mesh(box, mask_boundary=dirichlet)
        self.mesh = m
        self.E, self.Np = box.E, box.Np
        self.offset = self.E * self.Np
        self.lambda1 = 0.0 if bpmode else lambda1
        # singular all-Neumann problem (setup.cpp:210-214) that carries no mask
        self.all_neumann = self.lambda1 == 0.0 and not dirichlet
        self.bk = int(bool(bpmode))
        self.lam = np.concatenate([np.ones(self.offset), np.full(self.offset, self.lambda1)])
        self.ogs, self.off, self.gid, self.invd = ogs_setup(m["ids"])
        self.maskIds = np.nonzero(m["ids"] == 0)[0].astype(np.int32) if dirichlet else np.zeros(0, np.int32)
        self.struct = BPStruct(self.N, Nf, self.bk, self.E, self.offset, m["ggeo"].ctypes.data,
                               m["D"].ctypes.data, self.lam.ctypes.data, C.addressof(self.ogs),
                               self.invd.ctypes.data, self.maskIds.size,
                               self.maskIds.ctypes.data if self.maskIds.size else None)
        r = np.zeros(Nf * self.offset)
        if rhs_kind == 1:
            JW = m["ggeo"].reshape(self.E, 7, self.Np)[:, 0].ravel()
            v = (3.0 * np.pi * np.pi + self.lambda1) * JW * np.cos(np.pi * m["x"]) * np.cos(np.pi * m["y"]) * \
                np.cos(np.pi * m["z"])
            for f in range(Nf):
                r[f * self.offset:(f + 1) * self.offset] = v
        else:
            for f in range(Nf):
                r[f * self.offset:(f + 1) * self.offset] = urand(box.seed, 3 + f, self.offset)
        for f in range(Nf):
            r[self.maskIds + f * self.offset] = 0.0
        gs_many(self.ogs, r, Nf, self.offset)
        self.rhs = r

    def apply(self, q):
        Aq = np.zeros_like(q)
        lib().nbo_bp_apply(C.byref(self.struct), p(q), p(Aq))
        return Aq

    def wdot(self, x, y):
        f = lib().nbo_bp_wdot
        f.restype = C.c_double
        return f(C.byref(self.struct), p(x), p(y))

    def zero_mean(self, q):
        lib().nbo_bp_zero_mean(C.byref(self.struct), p(q))

    def pcg(self, r, x, tol, maxit, jacobi=True, flexible=False, fused=True):
        """solve() of nekBone/main.cpp:4-34: BPZeroMean around BPPCG when all-Neumann."""
        hr = np.zeros(maxit + 2)
        ha = np.zeros(maxit + 2)
        st = PCGStats()
        if self.all_neumann:
            self.zero_mean(r)
        lib().nbo_bp_pcg(C.byref(self.struct), int(jacobi), int(flexible), int(fused), p(r), p(x), C.c_double(tol), maxit,
                         p(hr), p(ha), C.byref(st))
        if self.all_neumann:
            self.zero_mean(x)
        return st, hr, ha

    def pcg_mixed(self, r, x, tol, maxit, prec=1, inner_tol=1e-5, max_cycles=20, jacobi=True, flexible=False):
        st = PCGStats()
        lib().nbo_bp_pcg_mixed(C.byref(self.struct), int(jacobi), int(flexible), int(prec), p(r), p(x),
                               C.c_double(tol), maxit, C.c_double(inner_tol), max_cycles, C.byref(st))
        return st

    def pcg_mixed_learned(self, r, x, tol, maxit, W, feat, prec=1, inner_tol=1e-5, max_cycles=20, window=20,
                          ratio=0.9):
        """Refinement with the learned preconditioner in the inner solves."""
        st = PCGStats()
        lib().nbo_bp_pcg_mixed_learned(C.byref(self.struct), p(np.ascontiguousarray(W, dtype=np.float32)), p(feat),
                                       window, C.c_double(ratio), int(prec), p(r), p(x), C.c_double(tol), maxit,
                                       C.c_double(inner_tol), max_cycles, C.byref(st))
        return st

    def invdiag(self):
        d = jacobi_local(self.N, self.E, self.offset, self.mesh["ggeo"], self.mesh["D"], self.lam)
        gs_many(self.ogs, d, 1, self.offset)
        return 1.0 / d

    def pcg_learned(self, r, x, tol, maxit, W, feat, window=20, ratio=0.9):
        hr = np.zeros(maxit + 2)
        ha = np.zeros(maxit + 2)
        st = PCGStats()
        lib().nbo_bp_pcg_learned(C.byref(self.struct), p(np.ascontiguousarray(W, dtype=np.float32)), p(feat), window,
                                 C.c_double(ratio), p(r), p(x), C.c_double(tol), maxit, p(hr), p(ha), C.byref(st))
        return st, hr, ha

    def __del__(self):
        try:
            ogs_free(self.ogs)
        except Exception:
            pass


def axhelm_bf16mma(bk, E, offset, ggeo, D, lam, q):
    """float32 arrays; ggeo and lam hold bf16 values"""
    Aq = np.zeros_like(q)
    lib().nbo_axhelm_bf16mma(int(bk), C.c_int32(E), C.c_int32(offset), p(ggeo), p(D), p(lam), p(q), p(Aq))
    return Aq


def axhelm_f16mma(bk, E, offset, ggeo, D, lam, q):
    """float32 arrays (FP32 ggeo and lam); fp16 matrix-core operands"""
    Aq = np.zeros_like(q)
    lib().nbo_axhelm_f16mma(int(bk), C.c_int32(E), C.c_int32(offset), p(ggeo), p(D), p(lam), p(q), p(Aq))
    return Aq


def axhelm_mma_n(f16, N, Nf, bk, E, offset, ggeo, D, lam, q):
    """float32 arrays, any order N and Nf fields (field f of q and Aq at
    f*offset); bf16 (f16=False: ggeo and lam hold bf16 values) or fp16
    matrix-core operands"""
    Aq = np.zeros_like(q)
    lib().nbo_axhelm_mma_n(int(bool(f16)), int(N), int(Nf), int(bk), C.c_int32(E), C.c_int32(offset), p(ggeo), p(D),
                           p(lam), p(q), p(Aq))
    return Aq


def f16_round(x):
    """float32 -> nearest-even IEEE binary16 value (as float32), by the C restatement."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    lib().nbo_f16_round_array(C.c_int64(x.size), p(x), p(y))
    return y


def sumsq_inv_f(a):
    f = lib().nbo_sumsq_inv_f
    f.restype = C.c_double
    return f(C.c_int64(a.size), p(np.ascontiguousarray(a, dtype=np.float32)))


def net_features(E, Np, ggeo):
    feat = np.empty(E * 7 * Np, dtype=np.float32)
    lib().nbo_net_features(C.c_int32(E), Np, p(ggeo), p(feat))
    return feat


def net_apply(E, W, feat, r, invDiag, w=None):
    zl = np.zeros(E * 512)
    lib().nbo_net_apply(C.c_int32(E), p(np.ascontiguousarray(W, dtype=np.float32)), p(feat), p(r), p(invDiag), p(w),
                        p(zl))
    return zl


def net_apply_mma(E, W, feat, r, invDiag, w=None):
    zl = np.zeros(E * 512)
    lib().nbo_net_apply_mma(C.c_int32(E), p(np.ascontiguousarray(W, dtype=np.float32)), p(feat), p(r), p(invDiag),
                            p(w), p(zl))
    return zl


def net_gradient(samples, chunk, pairStride, R, Ev, invNorm, W, feat, invDiag, G=None):
    """G given: gradient of an external loss whose derivative with respect to
    the unweighted element output invDiag (r + s y) is G (local vector)."""
    s = np.ascontiguousarray(samples, dtype=np.int32).reshape(-1, 2)
    g = np.zeros(3041, dtype=np.float32)
    loss = C.c_double(0.0)
    lib().nbo_net_gradient(C.c_int32(s.shape[0]), C.c_int32(chunk), p(s), C.c_int64(pairStride), p(R), p(Ev),
                           p(invNorm), p(np.ascontiguousarray(W, dtype=np.float32)), p(feat), p(invDiag), p(G),
                           p(g), C.byref(loss))
    return g, loss.value


def bf16_round(x):
    """float32 -> nearest-even bf16 value (as float32), by the C restatement
    (nbo_bf16r: the one bf16 rounding every oracle path uses)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    lib().nbo_bf16_round_array(C.c_int64(x.size), p(x), p(y))
    return y


class MGStruct(C.Structure):
    _fields_ = [("nlev", C.c_int), ("bps", C.c_void_p), ("invD", C.c_void_p), ("J", C.c_void_p),
                ("degree", C.c_int), ("c1", C.c_void_p), ("c2", C.c_void_p), ("coarse_iters", C.c_int),
                ("prec", C.c_int), ("ggeo32", C.c_void_p), ("lam32", C.c_void_p), ("D32", C.c_void_p),
                ("invD32", C.c_void_p), ("J32", C.c_void_p), ("W", C.c_void_p), ("feat", C.c_void_p)]


class OracleMG:
    """The oracle's p-multigrid preconditioner, built as nbx_bp_mg_setup
    builds it: one OracleBP per order, Jacobi diagonals, interpolation
    matrices, power-iteration lambda_max and Chebyshev coefficients."""

    def __init__(self, box: Box, orders, degree=2, coarse_iters=20, lo=0.1, hi=1.1, bpmode=True, lambda1=0.0,
                 dirichlet=True, power_iters=20, prec=0, Nf=1):
        assert orders[0] == box.N
        self.levels = []
        for N in orders:
            b = Box.make(N, box.NX, box.NY, box.NZ, box.periodic, box.perturb, box.zaspect, box.seed,
                         (box.x0, box.x1, box.y0, box.y1, box.z0, box.z1))
            self.levels.append(OracleBP(b, Nf=Nf, bpmode=bpmode, lambda1=lambda1, dirichlet=dirichlet))
        n = len(orders)
        self.fine = self.levels[0]
        self.bps = (BPStruct * n)(*[lv.struct for lv in self.levels])
        self.invD = [lv.invdiag() for lv in self.levels]
        self.invD_ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in self.invD])
        self.J = []
        for l in range(n - 1):
            Jm = np.zeros((orders[l] + 1) * (orders[l + 1] + 1))
            lib().nbo_mg_interp(orders[l], orders[l + 1], p(Jm))
            self.J.append(Jm)
        self.J_ptrs = (C.c_void_p * n)(*([a.ctypes.data for a in self.J] + [None]))
        power = lib().nbo_mg_power
        power.restype = C.c_double
        self.lmax = []
        self.c1 = np.zeros(n * degree)
        self.c2 = np.zeros(n * degree)
        for l in range(n - 1):
            lam = power(C.byref(self.bps, l * C.sizeof(BPStruct)), p(self.invD[l]), C.c_uint64(box.seed),
                        C.c_uint64(17 + l), power_iters)
            self.lmax.append(lam)
            c1 = np.zeros(degree)
            c2 = np.zeros(degree)
            lib().nbo_mg_cheb(C.c_double(lam), C.c_double(lo), C.c_double(hi), degree, p(c1), p(c2))
            self.c1[l * degree:(l + 1) * degree] = c1
            self.c2[l * degree:(l + 1) * degree] = c2
        # FP32 / bf16-geometry cycle: float copies (double -> float, bf16 by RNE of the float)
        def lo_vals(a):
            f = np.ascontiguousarray(a, dtype=np.float32)
            return bf16_round(f) if prec in (2, 3) else f
        self.f32 = []
        for l, lv in enumerate(self.levels):
            self.f32.append((lo_vals(lv.mesh["ggeo"]), lo_vals(lv.lam), np.ascontiguousarray(lv.mesh["D"], np.float32),
                             self.invD[l].astype(np.float32),
                             self.J[l].astype(np.float32) if l < n - 1 else np.zeros(1, np.float32)))
        arr = lambda i: (C.c_void_p * n)(*[t[i].ctypes.data for t in self.f32])
        self.f32_ptrs = [arr(i) for i in range(5)]
        self.struct = MGStruct(n, C.addressof(self.bps), C.addressof(self.invD_ptrs), C.addressof(self.J_ptrs),
                               degree, self.c1.ctypes.data, self.c2.ctypes.data, coarse_iters, prec,
                               *[C.addressof(a) for a in self.f32_ptrs], None, None)
        self._net = None

    def set_net(self, W=None):
        """Learned step on the fine level (library: nbx_bp_mg_set_learned) with
        network weights W (bf16 values as float32; None turns it off); the
        features are the fine mesh's normalised geometric factors."""
        if W is None:
            self.struct.W = self.struct.feat = None
            self._net = None
            return
        fine = self.fine
        feat = net_features(fine.E, fine.Np, fine.mesh["ggeo"])
        W = np.ascontiguousarray(W, dtype=np.float32)
        self._net = (W, feat)
        self.struct.W = W.ctypes.data
        self.struct.feat = feat.ctypes.data

    def vcycle(self, r):
        x = np.zeros_like(r)
        lib().nbo_mg_vcycle(C.byref(self.struct), 0, p(np.ascontiguousarray(r)), p(x))
        return x

    def pcg(self, r, x, tol, maxit):
        hr = np.zeros(maxit + 2)
        ha = np.zeros(maxit + 2)
        st = PCGStats()
        lib().nbo_bp_pcg_mg(C.byref(self.struct), p(r), p(x), C.c_double(tol), maxit, p(hr), p(ha), C.byref(st))
        return st, hr, ha
