"""Generate golden vectors from the reference's own CPU kernels.

Runs in the build container only (needs oracle/_ref, compiled by
`make -C oracle ref` from the nekBench tree).  Writes tests/golden/*.npz:
inputs and the outputs of
  axhelm/kernel/CPU/axhelm.c        axhelm_v0, axhelm_bk_v0, axhelm_n3_v0, axhelm_bk_n3_v0
  axhelm/axhelmReference.cpp        axhelmReference (lambda = 0 and 1.1)
  nekBone/kernel/CPU/weightedInnerProduct.c  weightedInnerProductUpdate (1 thread)
Inputs: a small perturbed box mesh's ggeo and GLL D from the oracle, q and
lambda from the splitmix64 streams.  The tests then pin the oracle (and the
GPU path through it) against these files on machines without the reference.
"""
import os
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
os.environ.setdefault("OMP_NUM_THREADS", "1")
import nbo  # noqa: E402

OUT = Path(__file__).resolve().parent


def case(N, NX, NY, NZ, seed):
    box = nbo.Box.make(N, NX, NY, NZ, perturb=0.1, seed=seed)
    m = nbo.box_mesh(box)
    E, Np = box.E, box.Np
    offset = E * Np
    q = nbo.urand(seed, 2, 3 * offset)
    lam = np.empty(2 * offset)
    lam[:offset] = 1.0 + 0.5 * nbo.urand(seed, 4, offset)     # lambda0 > 0
    lam[offset:] = 0.6 + 0.5 * nbo.urand(seed, 5, offset)     # lambda1 > 0
    out = dict(N=N, E=E, offset=offset, ggeo=m["ggeo"], D=m["D"], q=q, lam=lam)
    for k in ("axhelm_v0", "axhelm_bk_v0"):
        out[k] = nbo.ref_axhelm(N, k, E, offset, m["ggeo"], m["D"], lam, q[:offset].copy())
    for k in ("axhelm_n3_v0", "axhelm_bk_n3_v0"):
        out[k] = nbo.ref_axhelm(N, k, E, offset, m["ggeo"], m["D"], lam, q.copy())
    for l1 in (0.0, 1.1):
        out[f"axhelmReference_{l1}"] = nbo.ref_axhelm_reference(N + 1, E, l1, m["ggeo"], m["D"], q[:offset].copy())
    w = 1.0 / (1.0 + np.floor(4.0 * np.abs(nbo.urand(seed, 6, offset))))  # 1, 1/2, 1/3, 1/4, 1/5
    a, b = q[:offset].copy(), nbo.urand(seed, 7, offset)
    out["wip_w"], out["wip_a"], out["wip_b"] = w, a, b
    out["weightedInnerProductUpdate"] = nbo.ref_weighted_inner_product_update(1, offset, offset, w, a, b)
    return out


def main():
    if not nbo.ref_available(7):
        raise SystemExit("oracle/_ref not built: make -C oracle ref")
    for (N, NX, NY, NZ, seed) in [(2, 2, 2, 1, 11), (7, 3, 1, 1, 12), (9, 2, 1, 1, 13)]:
        d = case(N, NX, NY, NZ, seed)
        path = OUT / f"axhelm_N{N}.npz"
        np.savez_compressed(path, **d)
        print(f"wrote {path.name}: E={d['E']} ({path.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
