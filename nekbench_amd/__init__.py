"""nekbench_amd -- MI355X (gfx950) nekBench solver hot path.

Python view of the C ABI declared in include/nbx.h (libnbx.so, built in-tree
by nekbench_amd/build.py).  Device buffers are passed as torch tensors on a
ROCm device (only their data pointers cross the boundary); host buffers as
numpy arrays.  There is no fallback: if libnbx.so is missing or a call fails,
an exception is raised.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libnbx.so"

NBX_F64, NBX_F32, NBX_BF16, NBX_F16, NBX_BF16_MFMA = 0, 1, 2, 3, 4
NBX_DOUBLE, NBX_FLOAT, NBX_INT, NBX_LONG = 0, 1, 2, 3
NBX_ADD, NBX_MUL, NBX_MIN, NBX_MAX = 0, 1, 2, 3
NBX_SUCCESS, NBX_ERR_INVALID, NBX_ERR_UNSUPPORTED, NBX_ERR_HIP, NBX_ERR_ALLOC, NBX_ERR_DIVERGED = 0, -1, -2, -3, -4, -5
NBX_WDOT_NOALLOC_MAX_N = 134217728  # include/nbx.h
PRECON_COPY, PRECON_JACOBI, PRECON_LEARNED, PRECON_MG = 0, 1, 2, 3
# reference ggeo factor order (core/params.h:55-62)
GWJID, G00ID, G01ID, G02ID, G11ID, G12ID, G22ID, NGGEO = 0, 1, 2, 3, 4, 5, 6, 7

_lib = None


class NbxError(RuntimeError):
    pass


class Box(C.Structure):
    """nbx_box: box mesh description (include/nbx.h)."""
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
    def Nelements(self):
        return self.NX * self.NY * self.NZ

    @property
    def Np(self):
        return (self.N + 1) ** 3


class BPOptions(C.Structure):
    """nbx_bp_options (include/nbx.h)."""
    _fields_ = [("box", Box), ("Nfields", C.c_int), ("bpmode", C.c_int), ("lambda1", C.c_double),
                ("dirichlet", C.c_int), ("precon", C.c_int), ("flexible", C.c_int), ("verbose", C.c_int),
                ("rhs_kind", C.c_int), ("fused", C.c_int)]


class PCGStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("rdotr0", C.c_double),
                ("rdotr", C.c_double), ("elapsed_s", C.c_double), ("cycles", C.c_int),
                ("fallback_iteration", C.c_int)]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_SIGS = {
    "nbx_last_error": (C.c_char_p, []),
    "nbx_version": (_I, []),
    "nbx_device_synchronize": (_I, []),
    "nbx_axhelm": (_I, [_I, _I, _I, _I, C.c_int32, C.c_int32, _P, _P, _P, _P, _P, _P, _P]),
    "nbx_axhelm_by_name": (_I, [C.c_char_p, _I, C.c_int32, C.c_int32, _P, _P, _P, _P, _P, _P, _P]),
    "nbx_update_jacobi_local": (_I, [_I, C.c_int32, C.c_int32, _P, _P, _P, _P, _P, _P]),
    "nbx_ogs_setup": (_P, [C.c_int32, _P, _I]),
    "nbx_ogs_gather_scatter_many": (_I, [_P, _P, _I, C.c_int32, _I, _I, _P]),
    "nbx_ogs_gather_scatter": (_I, [_P, _P, _I, _I, _P]),
    "nbx_ogs_inv_degree": (_P, [_P]),
    "nbx_ogs_sizes": (_I, [_P, _P, _P, _P, _P, _P]),
    "nbx_ogs_copy_csr": (_I, [_P, _P, _P]),
    "nbx_ogs_free": (None, [_P]),
    "nbx_mask": (_I, [C.c_int32, _P, _P, _P]),
    "nbx_gll": (_I, [_I, _P, _P, _P]),
    "nbx_urand_fill": (_I, [C.c_uint64, C.c_uint64, C.c_int64, _P]),
    "nbx_box_mesh": (_I, [C.POINTER(Box), _I, _P, _P, _P, _P, _P, _P]),
    "nbx_weighted_inner_product": (_I, [C.c_int32, C.c_int32, _I, _P, _P, _P, _P, _P]),
    "nbx_weighted_inner_product_noalloc": (_I, [C.c_int32, C.c_int32, _I, _P, _P, _P, _P, _P]),
    "nbx_scaled_add": (_I, [C.c_int32, _D, _P, _D, _P, _P]),
    "nbx_box_mesh_device": (_I, [C.POINTER(Box), _I, _P, _P, _P]),
    "nbx_ogs_setup_device": (_P, [C.c_int32, _P, _I]),
    "nbx_triad": (_I, [C.c_int32, _D, _P, _P, _P, _P]),
    "nbx_copy": (_I, [C.c_int32, _P, _P, _P]),
    "nbx_read_sum": (_I, [C.c_int32, _P, _P, _P]),
    "nbx_lds_probe": (_I, [C.c_int32, _P, _P]),
    "nbx_bp_setup": (_P, [C.POINTER(BPOptions), _P]),
    "nbx_bp_free": (None, [_P]),
    "nbx_bp_sizes": (_I, [_P, _P, _P, _P, _P]),
    "nbx_bp_device_array": (_P, [_P, C.c_char_p]),
    "nbx_bp_ogs": (_P, [_P]),
    "nbx_bp_ax_operator": (_I, [_P, _P, _P, _P, _P]),
    "nbx_bp_pcg": (_I, [_P, _P, _P, _P, _D, _I, C.POINTER(PCGStats)]),
    "nbx_bp_history": (_I, [_P, _P, _P, _I]),
    "nbx_bp_pcg_mixed": (_I, [_P, _P, _P, _P, _D, _I, _I, _D, _I, C.POINTER(PCGStats)]),
}


def lib() -> C.CDLL:
    """Load libnbx.so (built by nekbench_amd.build); raises if it is absent."""
    global _lib
    if _lib is None:
        # torch ships its own libamdhip64.so.7; loading it first makes libnbx's
        # DT_NEEDED bind to that same runtime, so device pointers and streams
        # from torch are valid inside the library (one HIP runtime per process)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not LIB_PATH.exists():
            raise NbxError(f"{LIB_PATH} not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
        l = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def last_error() -> str:
    return lib().nbx_last_error().decode()


def check(rc: int, what: str = "nbx call") -> None:
    if rc != 0:
        raise NbxError(f"{what} failed ({rc}): {last_error()}")


def ptr(t) -> int | None:
    """Device (torch) or host (numpy) data pointer; None stays NULL."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        assert t.flags["C_CONTIGUOUS"]
        return t.ctypes.data
    assert t.is_contiguous()
    return t.data_ptr()


def stream_ptr(stream=None) -> int | None:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


# ---------------------------------------------------------------- host setup
def gll(N: int):
    z = np.zeros(N + 1)
    w = np.zeros(N + 1)
    D = np.zeros((N + 1) * (N + 1))
    check(lib().nbx_gll(N, ptr(z), ptr(w), ptr(D)), "nbx_gll")
    return z, w, D


def urand(seed: int, stream: int, n: int) -> np.ndarray:
    out = np.empty(n)
    check(lib().nbx_urand_fill(seed, stream, n, ptr(out)), "nbx_urand_fill")
    return out


def box_mesh(box: Box, mask_boundary: bool = False, nodes: bool = False, ids: bool = True):
    """Host mesh inputs: dict with ggeo [E*7*Np], ids [E*Np] (int64), x,y,z."""
    E, Np = box.Nelements, box.Np
    out = {"ggeo": np.empty(E * NGGEO * Np)}
    if nodes:
        for c in "xyz":
            out[c] = np.empty(E * Np)
    if ids:
        out["ids"] = np.empty(E * Np, dtype=np.int64)
    minJ = C.c_double(0.0)
    check(lib().nbx_box_mesh(C.byref(box), int(mask_boundary), ptr(out.get("x")), ptr(out.get("y")),
                             ptr(out.get("z")), ptr(out["ggeo"]), ptr(out.get("ids")), C.byref(minJ)),
          "nbx_box_mesh")
    out["minJ"] = minJ.value
    return out


# ---------------------------------------------------------------- device ops
def axhelm(N, Nfields, bk, Nelements, offset, ggeo, D, lam, q, Aq, elementList=None, prec=NBX_F64, stream=None):
    check(lib().nbx_axhelm(N, Nfields, int(bk), prec, Nelements, offset, ptr(elementList), ptr(ggeo), ptr(D),
                           ptr(lam), ptr(q), ptr(Aq), stream_ptr(stream)), "nbx_axhelm")


def axhelm_by_name(name, N, Nelements, offset, ggeo, D, lam, q, Aq, elementList=None, stream=None):
    check(lib().nbx_axhelm_by_name(name.encode(), N, Nelements, offset, ptr(elementList), ptr(ggeo), ptr(D),
                                   ptr(lam), ptr(q), ptr(Aq), stream_ptr(stream)), f"nbx_axhelm_by_name({name})")


def update_jacobi_local(N, Nelements, offset, ggeo, D, lam, diag, mapB=None, stream=None):
    check(lib().nbx_update_jacobi_local(N, Nelements, offset, ptr(mapB), ptr(ggeo), ptr(D), ptr(lam), ptr(diag),
                                        stream_ptr(stream)), "nbx_update_jacobi_local")


def mask(maskIds, q, stream=None):
    check(lib().nbx_mask(maskIds.numel(), ptr(maskIds), ptr(q), stream_ptr(stream)), "nbx_mask")


def weighted_inner_product(N, offset, Nfields, w, x, y, stream=None) -> float:
    out = C.c_double(0.0)
    check(lib().nbx_weighted_inner_product(N, offset, Nfields, ptr(w), ptr(x), ptr(y), C.byref(out),
                                           stream_ptr(stream)), "nbx_weighted_inner_product")
    return out.value


def weighted_inner_product_noalloc(N, offset, Nfields, w, x, y, out, stream=None) -> None:
    """<x,y>_w into out (device tensor, one double) on `stream`, no host wait;
    the library's fixed device array holds the partials (no allocation)."""
    check(lib().nbx_weighted_inner_product_noalloc(N, offset, Nfields, ptr(w), ptr(x), ptr(y), ptr(out),
                                                   stream_ptr(stream)), "nbx_weighted_inner_product_noalloc")


def scaled_add(N, alpha, x, beta, y, stream=None):
    check(lib().nbx_scaled_add(N, alpha, ptr(x), beta, ptr(y), stream_ptr(stream)), "nbx_scaled_add")


class Ogs:
    """Gather-scatter handle (nbx_ogs_setup / nbx_ogs_gather_scatter_many)."""

    _types = {"double": NBX_DOUBLE, "float": NBX_FLOAT, "int": NBX_INT, "long long int": NBX_LONG}
    _ops = {"add": NBX_ADD, "mul": NBX_MUL, "min": NBX_MIN, "max": NBX_MAX}

    def __init__(self, ids: np.ndarray, verbose: bool = False, handle=None):
        self._owned = handle is None
        if handle is None:
            ids = np.ascontiguousarray(ids, dtype=np.int64)
            handle = lib().nbx_ogs_setup(ids.size, ptr(ids), int(verbose))
            if not handle:
                raise NbxError(f"nbx_ogs_setup failed: {last_error()}")
        self.h = handle

    def sizes(self):
        v = [C.c_int32(0) for _ in range(5)]
        check(lib().nbx_ogs_sizes(self.h, *[C.byref(x) for x in v]), "nbx_ogs_sizes")
        return dict(zip(["N", "Nlocal", "NlocalGather", "NsharedGather", "Nshared"], [x.value for x in v]))

    def csr(self):
        s = self.sizes()
        off = np.zeros(s["NlocalGather"] + 1, dtype=np.int32)
        ids = np.zeros(max(s["Nlocal"], 1), dtype=np.int32)
        check(lib().nbx_ogs_copy_csr(self.h, ptr(off), ptr(ids)), "nbx_ogs_copy_csr")
        return off, ids[: s["Nlocal"]]

    def gather_scatter_many(self, v, k, stride, type_="double", op="add", stream=None):
        check(lib().nbx_ogs_gather_scatter_many(self.h, ptr(v), k, stride, self._types[type_], self._ops[op],
                                                stream_ptr(stream)), "nbx_ogs_gather_scatter_many")

    def gather_scatter(self, v, type_="double", op="add", stream=None):
        check(lib().nbx_ogs_gather_scatter(self.h, ptr(v), self._types[type_], self._ops[op], stream_ptr(stream)),
              "nbx_ogs_gather_scatter")

    def inv_degree_ptr(self) -> int:
        return lib().nbx_ogs_inv_degree(self.h)

    def free(self):
        if self.h and self._owned:
            lib().nbx_ogs_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BP:
    """BP driver handle: nekBone's setup + BPPCG + BPAxOperator on one GPU
    (nbx_bp_setup / nbx_bp_pcg / nbx_bp_ax_operator)."""

    def __init__(self, box: Box, Nfields=1, bpmode=False, lambda1=0.0, dirichlet=True, precon=PRECON_JACOBI,
                 flexible=False, verbose=False, rhs_kind=0, fused=True, stream=None):
        self.opt = BPOptions(box, Nfields, int(bpmode), lambda1, int(dirichlet), precon, int(flexible),
                             int(verbose), rhs_kind, int(fused))
        self.h = lib().nbx_bp_setup(C.byref(self.opt), stream_ptr(stream))
        if not self.h:
            raise NbxError(f"nbx_bp_setup failed: {last_error()}")
        v = [C.c_int32(0) for _ in range(4)]
        check(lib().nbx_bp_sizes(self.h, *[C.byref(x) for x in v]), "nbx_bp_sizes")
        self.Nelements, self.Np, self.fieldOffset, self.Nmasked = [x.value for x in v]
        self.Nfields = Nfields
        self.Nlocal = self.Nelements * self.Np

    def device_array(self, name, n, dtype=None):
        """torch view (no copy) of a device array owned by the handle."""
        import torch
        addr = lib().nbx_bp_device_array(self.h, name.encode())
        if not addr:
            raise NbxError(last_error())
        dtype = dtype or (torch.int32 if name == "maskIds" else torch.float64)
        return _wrap_device(addr, n, dtype)

    def ogs(self) -> Ogs:
        return Ogs(None, handle=lib().nbx_bp_ogs(self.h))

    def ax_operator(self, q, Aq, lam=None) -> float:
        pAp = C.c_double(0.0)
        check(lib().nbx_bp_ax_operator(self.h, ptr(lam), ptr(q), ptr(Aq), C.byref(pAp)), "nbx_bp_ax_operator")
        return pAp.value

    def pcg(self, r, x, tol, maxit, lam=None) -> PCGStats:
        st = PCGStats()
        rc = lib().nbx_bp_pcg(self.h, ptr(lam), ptr(r), ptr(x), tol, maxit, C.byref(st))
        if rc < 0:
            raise NbxError(f"nbx_bp_pcg failed ({rc}): {last_error()}")
        return st

    def pcg_mixed(self, r, x, tol, maxit, prec=NBX_F32, inner_tol=1e-5, max_cycles=20, lam=None) -> PCGStats:
        st = PCGStats()
        rc = lib().nbx_bp_pcg_mixed(self.h, ptr(lam), ptr(r), ptr(x), tol, maxit, prec, inner_tol, max_cycles,
                                    C.byref(st))
        if rc < 0:
            raise NbxError(f"nbx_bp_pcg_mixed failed ({rc}): {last_error()}")
        return st

    def mg_setup(self, orders=(7, 3, 1), degree=2, coarse_iters=20, lo=0.1, hi=1.1, prec=NBX_F64):
        o = np.ascontiguousarray(orders, dtype=np.int32)
        check(lib().nbx_bp_mg_setup(self.h, o.size, ptr(o), degree, coarse_iters, lo, hi, prec),
              "nbx_bp_mg_setup")

    def mg_info(self, level):
        N = C.c_int(0)
        lam = C.c_double(0.0)
        rc = lib().nbx_bp_mg_info(self.h, level, C.byref(N), C.byref(lam))
        if rc < 0:
            raise NbxError(f"nbx_bp_mg_info failed ({rc}): {last_error()}")
        return N.value, lam.value

    def mg_use_graph(self, on=True):
        check(lib().nbx_bp_mg_use_graph(self.h, int(on)), "nbx_bp_mg_use_graph")

    def mg_apply(self, r, z):
        check(lib().nbx_bp_mg_apply(self.h, ptr(r), ptr(z)), "nbx_bp_mg_apply")

    def mg_set_learned(self, on=True):
        """The fine level ends each V-cycle with the attached network's correction."""
        check(lib().nbx_bp_mg_set_learned(self.h, int(on)), "nbx_bp_mg_set_learned")

    def mg_learned(self) -> bool:
        return bool(lib().nbx_bp_mg_get_learned(self.h))

    def make_rhs(self, seed, out):
        check(lib().nbx_bp_make_rhs(self.h, seed, ptr(out)), "nbx_bp_make_rhs")

    def zero_mean(self, q):
        check(lib().nbx_bp_zero_mean(self.h, ptr(q)), "nbx_bp_zero_mean")

    @property
    def all_neumann(self) -> bool:
        return bool(lib().nbx_bp_all_neumann(self.h))

    def history(self, n):
        rd = np.zeros(n)
        al = np.zeros(n)
        k = lib().nbx_bp_history(self.h, ptr(rd), ptr(al), n)
        if k < 0:
            raise NbxError(last_error())
        return rd[:k], al[:k]

    def free(self):
        if getattr(self, "h", None):
            lib().nbx_bp_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _wrap_device(addr: int, n: int, dtype):
    """Torch tensor over device memory owned by a library handle: a
    non-owning view through __cuda_array_interface__, or, if this torch build
    cannot import one, a device-to-device copy (hipMemcpy)."""
    import torch
    typestr = {torch.float64: "<f8", torch.int32: "<i4", torch.float32: "<f4"}[dtype]

    class _Holder:
        __cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (addr, False), "version": 3,
                                    "strides": None}
    try:
        t = torch.as_tensor(_Holder(), device="cuda")
        if t.data_ptr() == addr and t.dtype == dtype and t.numel() == n:
            return t
    except Exception:
        pass
    out = torch.empty(n, dtype=dtype, device="cuda")
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    torch.cuda.synchronize()
    rc = hip.hipMemcpy(out.data_ptr(), addr, n * out.element_size(), 3)  # hipMemcpyDeviceToDevice
    if rc != 0:
        raise NbxError(f"hipMemcpy failed ({rc})")
    return out


_SIGS.update({
    "nbx_net_num_params": (_I, []),
    "nbx_bp_zero_mean": (_I, [_P, _P]),
    "nbx_bp_make_rhs": (_I, [_P, C.c_uint64, _P]),
    "nbx_bp_mg_setup": (_I, [_P, _I, _P, _I, _I, _D, _D, _I]),
    "nbx_bp_mg_info": (_I, [_P, _I, _P, _P]),
    "nbx_bp_mg_apply": (_I, [_P, _P, _P]),
    "nbx_bp_mg_use_graph": (_I, [_P, _I]),
    "nbx_bp_mg_set_learned": (_I, [_P, _I]),
    "nbx_bp_mg_get_learned": (_I, [_P]),
    "nbx_bp_all_neumann": (_I, [_P]),
    "nbx_net_create": (_P, [C.c_uint64, _P]),
    "nbx_net_free": (None, [_P]),
    "nbx_net_get_params": (_I, [_P, _P]),
    "nbx_net_set_params": (_I, [_P, _P]),
    "nbx_net_save": (_I, [_P, C.c_char_p]),
    "nbx_net_load": (_P, [C.c_char_p, _P]),
    "nbx_bp_attach_net": (_I, [_P, _P]),
    "nbx_bp_set_fallback": (_I, [_P, _I, _D]),
    "nbx_bp_update_jacobi": (_I, [_P]),
    "nbx_net_apply": (_I, [_P, _P, _P, _I]),
    "nbx_net_generate_pairs": (_I, [_P, _P, _I, C.c_uint64, _P, _I, _D]),
    "nbx_net_generate_mg_pairs": (_I, [_P, _P, _I, C.c_uint64, _I, _I, _D]),
    "nbx_net_copy_pairs": (_I, [_P, _I, _I, _P, _P]),
    "nbx_net_set_pairs": (_I, [_P, _P, _I, _P, _P]),
    "nbx_net_gradient": (_I, [_P, _P, _P, _I, _I, _P, _P, _P]),
    "nbx_net_train": (_I, [_P, _P, _I, _I, C.c_float, C.c_uint64, _P]),
    "nbx_net_train_energy": (_I, [_P, _P, _I, C.c_float, C.c_uint64, _I, _P]),
    "nbx_net_set_inference": (_I, [_P, _I]),
})


class Net:
    """Learned element preconditioner (nbx_net_*)."""

    def __init__(self, seed=1, handle=None, stream=None):
        l = lib()
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name, None)
            if f is not None:
                f.restype, f.argtypes = res, args
        self.h = handle or l.nbx_net_create(seed, stream_ptr(stream))
        if not self.h:
            raise NbxError(f"nbx_net_create failed: {last_error()}")
        self.nparams = l.nbx_net_num_params()
        self.npairs = 0

    @classmethod
    def load(cls, path, stream=None):
        l = lib()
        l.nbx_net_load.restype, l.nbx_net_load.argtypes = _P, [C.c_char_p, _P]
        h = l.nbx_net_load(str(path).encode(), stream_ptr(stream))
        if not h:
            raise NbxError(f"nbx_net_load failed: {last_error()}")
        return cls(handle=h)

    def save(self, path):
        check(lib().nbx_net_save(self.h, str(path).encode()), "nbx_net_save")

    def params(self) -> np.ndarray:
        out = np.zeros(self.nparams, dtype=np.float32)
        check(lib().nbx_net_get_params(self.h, ptr(out)), "nbx_net_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().nbx_net_set_params(self.h, ptr(p)), "nbx_net_set_params")

    def attach(self, bp: "BP"):
        check(lib().nbx_bp_attach_net(bp.h, self.h), "nbx_bp_attach_net")
        bp._net = self  # the handle uses the weights: keep them alive as long as it

    def set_inference(self, mfma: bool):
        check(lib().nbx_net_set_inference(self.h, int(bool(mfma))), "nbx_net_set_inference")

    def apply(self, bp: "BP", r, zl, weighted=True):
        check(lib().nbx_net_apply(bp.h, ptr(r), ptr(zl), int(weighted)), "nbx_net_apply")

    def generate_pairs(self, bp: "BP", nseeds, seed0, snapshots, tol_exact=1e-11) -> int:
        snaps = np.ascontiguousarray(snapshots, dtype=np.int32)
        k = lib().nbx_net_generate_pairs(self.h, bp.h, nseeds, seed0, ptr(snaps), snaps.size, tol_exact)
        if k < 0:
            raise NbxError(f"nbx_net_generate_pairs failed: {last_error()}")
        self.npairs = k
        return k

    def generate_mg_pairs(self, bp: "BP", nseeds, seed0, kcount, kfirst=0, tol_exact=1e-11) -> int:
        """(residual, error) pairs left by one V-cycle on the iterates
        kfirst .. kfirst+kcount-1 of the multigrid PCG (training data of the
        learned multigrid step)."""
        k = lib().nbx_net_generate_mg_pairs(self.h, bp.h, nseeds, seed0, kfirst, kcount, tol_exact)
        if k < 0:
            raise NbxError(f"nbx_net_generate_mg_pairs failed: {last_error()}")
        self.npairs = k
        return k

    def pairs(self, bp: "BP", first=0, count=None):
        """Host copies (float32 arrays, one row per pair) of the stored pairs."""
        if count is None:
            count = self.npairs - first
        R = np.zeros((count, bp.Nlocal), dtype=np.float32)
        E = np.zeros_like(R)
        check(lib().nbx_net_copy_pairs(self.h, first, count, ptr(R), ptr(E)), "nbx_net_copy_pairs")
        return R, E

    def set_pairs(self, bp: "BP", R, E):
        """R, E: device float32 tensors of shape (npairs, local nodes)."""
        assert R.dim() == 2 and R.shape == E.shape
        check(lib().nbx_net_set_pairs(self.h, bp.h, R.shape[0], ptr(R), ptr(E)), "nbx_net_set_pairs")
        self.npairs = R.shape[0]

    def gradient(self, bp: "BP", samples, chunk=16, G=None):
        s = np.ascontiguousarray(samples, dtype=np.int32).reshape(-1, 2)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_double(0.0)
        check(lib().nbx_net_gradient(self.h, bp.h, ptr(s), s.shape[0], chunk, ptr(G), ptr(g), C.byref(loss)),
              "nbx_net_gradient")
        return g, loss.value

    def train(self, bp: "BP", steps, batch=2048, lr=1e-3, seed=1):
        hist = np.zeros(max(steps, 1))
        check(lib().nbx_net_train(self.h, bp.h, steps, batch, C.c_float(lr), seed, ptr(hist)), "nbx_net_train")
        return hist[:steps]

    def train_energy(self, bp: "BP", steps, lr=1e-3, seed=1, mode=0):
        hist = np.zeros(max(steps, 1))
        check(lib().nbx_net_train_energy(self.h, bp.h, steps, C.c_float(lr), seed, mode, ptr(hist)),
              "nbx_net_train_energy")
        return hist[:steps]

    def free(self):
        if getattr(self, "h", None):
            lib().nbx_net_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _bp_set_fallback(self, window, ratio):
    l = lib()
    l.nbx_bp_set_fallback.restype, l.nbx_bp_set_fallback.argtypes = _I, [_P, _I, _D]
    check(l.nbx_bp_set_fallback(self.h, window, ratio), "nbx_bp_set_fallback")


def _bp_update_jacobi(self):
    l = lib()
    l.nbx_bp_update_jacobi.restype, l.nbx_bp_update_jacobi.argtypes = _I, [_P]
    check(l.nbx_bp_update_jacobi(self.h), "nbx_bp_update_jacobi")


BP.set_fallback = _bp_set_fallback
BP.update_jacobi = _bp_update_jacobi
