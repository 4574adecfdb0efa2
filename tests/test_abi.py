"""The C-ABI library loads and exports every symbol include/nbx.h declares;
host-side (no GPU) entry points agree with the oracle bit for bit."""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "nbx.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(nbx_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("nbx_axhelm", "nbx_ogs_setup", "nbx_ogs_gather_scatter_many", "nbx_bp_pcg",
                 "nbx_bp_ax_operator", "nbx_box_mesh"):
        assert must in names


def test_library_exports_every_declared_symbol(nbx):
    import ctypes
    lib = ctypes.CDLL(str(nbx.LIB_PATH))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    assert lib.nbx_version() >= 100


def test_library_has_gfx950_code_object(nbx):
    data = nbx.LIB_PATH.read_bytes()
    assert b"gfx950" in data


def test_host_gll_matches_oracle(nbx, nbo):
    for N in (1, 2, 3, 7, 9, 15):
        z, w, D = nbx.gll(N)
        oz, ow, oD = nbo.gll(N)
        assert np.array_equal(z, oz) and np.array_equal(w, ow) and np.array_equal(D, oD)


def test_host_urand_matches_oracle(nbx, nbo):
    for seed, stream in ((1, 2), (12345, 3), (2 ** 63 + 5, 1)):
        assert np.array_equal(nbx.urand(seed, stream, 1000), nbo.urand(seed, stream, 1000))


@pytest.mark.parametrize("kw", [dict(N=3, NX=3, NY=2, NZ=4, perturb=0.1, seed=5),
                                dict(N=7, NX=2, NY=2, NZ=3, perturb=0.1, zaspect=10.0, seed=6),
                                dict(N=2, NX=3, NY=3, NZ=3, periodic=1),
                                dict(N=9, NX=2, NY=1, NZ=1, perturb=0.0)])
@pytest.mark.parametrize("mask", [False, True])
def test_host_mesh_matches_oracle_bitwise(nbx, nbo, kw, mask):
    b = nbx.Box.make(**kw)
    ob = nbo.Box.make(**kw)
    got = nbx.box_mesh(b, mask_boundary=mask, nodes=True)
    ref = nbo.box_mesh(ob, mask_boundary=mask)
    for k in ("ggeo", "x", "y", "z", "ids"):
        assert np.array_equal(got[k], ref[k]), k
    assert got["minJ"] == ref["minJ"]


def test_host_errors_are_reported(nbx):
    bad = nbx.Box.make(0, 1, 1, 1)
    with pytest.raises(nbx.NbxError, match="bad sizes"):
        nbx.box_mesh(bad)


def test_native_symbol_libraries_export_the_reference_names(nbx):
    """nekbench_amd/native/libaxhelm_N<N>.so (N = 1..15) export every symbol
    include/nbx_native_axhelm.h declares -- the reference's native kernel
    names (axhelm/kernel/CPU/axhelm.c:27,113,229,307 and the Partial forms) --
    and report their order."""
    import ctypes
    text = (ROOT / "include" / "nbx_native_axhelm.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\b(axhelm[A-Za-z0-9_]*)\s*\(", text)))
    assert {"axhelm_v0", "axhelm_bk_v0", "axhelm_n3_v0", "axhelm_bk_n3_v0", "axhelmPartial_bk_v0"} <= set(names)
    for n in range(1, 16):
        lib = ctypes.CDLL(str(ROOT / "nekbench_amd" / "native" / f"libaxhelm_N{n}.so"))
        missing = [s for s in names if not hasattr(lib, s)]
        assert not missing, (n, missing)
        assert lib.axhelm_native_order() == n and lib.axhelm_native_status() == 0
