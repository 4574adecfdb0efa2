"""The allocation-free weighted inner product (nbx_weighted_inner_product_noalloc)
against the allocating entry point it varies (nbx_weighted_inner_product):
identical bits on two calls in a row, the second with a smaller input, so the
fixed device array's leftovers from the first call cannot leak into the
second; sizes past the array's maximum are refused before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_wdot_matches_allocating_entry_point(nbx, dev):
    import torch
    lib = nbx.lib()
    rng = np.random.default_rng(71)
    out = torch.full((1,), np.nan, dtype=torch.float64, device=dev)
    # (N, Nfields): a multi-tile three-field case, then a smaller one-field case
    for N, Nf in ((1_000_003, 3), (40_001, 1)):
        off = N + 17 if Nf > 1 else N
        w = torch.from_numpy(rng.uniform(0.1, 1.0, N)).to(dev)
        x = torch.from_numpy(rng.uniform(-1, 1, Nf * off)).to(dev)
        y = torch.from_numpy(rng.uniform(-1, 1, Nf * off)).to(dev)
        ref = ctypes.c_double(0.0)
        s = torch.cuda.current_stream().cuda_stream
        assert lib.nbx_weighted_inner_product(N, off, Nf, w.data_ptr(), x.data_ptr(), y.data_ptr(),
                                              ctypes.byref(ref), s) == 0, nbx.last_error()
        rc = lib.nbx_weighted_inner_product_noalloc(N, off, Nf, w.data_ptr(), x.data_ptr(), y.data_ptr(),
                                                    out.data_ptr(), s)
        assert rc == 0, nbx.last_error()
        torch.cuda.synchronize()
        got = float(out.cpu()[0])
        print(f"noalloc wdot N={N} Nf={Nf}: {got!r} vs {ref.value!r}", flush=True)
        assert np.float64(got).tobytes() == np.float64(ref.value).tobytes()


def test_noalloc_wdot_refuses_sizes_past_its_array(nbx, dev):
    import torch
    lib = nbx.lib()
    t = torch.zeros(4, dtype=torch.float64, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    n = nbx.NBX_WDOT_NOALLOC_MAX_N + 1  # refused on the host before any launch
    rc = lib.nbx_weighted_inner_product_noalloc(n, n, 1, t.data_ptr(), t.data_ptr(), t.data_ptr(), t.data_ptr(), s)
    assert rc == nbx.NBX_ERR_INVALID and "exceeds the maximum" in nbx.last_error()
