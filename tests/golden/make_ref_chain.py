"""Writes tests/golden/ref_chain_c1.npz: posterior means and Monte-Carlo
standard errors from a long chain of the reference's own CPU path
(oracle/_ref: updateZ, updateSpatCoef, updateSpatCoefPrec from
bsglmm/covar.cpp, compiled unchanged) on the C1 configuration:
32x32x32 grid, ellipsoid mask V = 7,056, N = 50, P = 3.

Semantics are the reference CPU path's (truncation at 0, covar.cpp:317,320;
Wishart df = TOTVOX + P, covar.cpp:291) in the GPU sweep order
(mcmc.cpp:311-394) without the WM coefficient (no WM image: b = 0; the CPU
branch's updateBeta lives in mcmc.cpp, which needs CUDA to build).

Saved: mean and MCSE (50 batch means) of every coefficient map in mask-rank
order, and of the lesion-probability map Phi(beta_0) (compute_prb,
covar.cpp:350-364, NSUBTYPES = 1). Run (a few minutes):

    make -C oracle && python tests/golden/make_ref_chain.py
"""
import ctypes as C
import os
import sys
import time

import numpy as np
import scipy.special as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402
from bsglmm_amd import synth  # noqa: E402

BURNIN, DRAWS, BATCHES = 1000, 10000, 50
SEED_DAT = (5118772087888418775, 289433066200759127, 1964621368)


def main():
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref is not built")
    d = synth.generate("c1")
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    covar = d["covar"]
    P, N = covar.shape
    V, VP = lay["totvox"], lay["totvoxp"]
    oracle.RefCPU(lay, covar, oracle.xtx_reference(covar))
    z = np.zeros((N, V), np.float32)
    data = np.ascontiguousarray(d["data"])
    wm = np.ones(V, np.float32)
    sc = np.zeros((P, VP), np.float32)
    prec = np.eye(P, dtype=np.float32)
    seed = oracle.seed_array(SEED_DAT)
    f, u8 = C.POINTER(C.c_float), C.POINTER(C.c_uint8)
    args = (z.ctypes.data_as(f), data.ctypes.data_as(u8), wm.ctypes.data_as(f), sc.ctypes.data_as(f),
            covar.ctypes.data_as(f), prec.ctypes.data_as(f), seed)
    hi, hs = lay["host_idx"], lay["host_idx_sc"]
    ids = np.flatnonzero(hi >= 0)
    col = np.empty(V, np.int64)
    col[hi[ids]] = hs[ids]  # mask rank -> dilated rank
    per = DRAWS // BATCHES
    bsum = np.zeros((BATCHES, P + 1, V))
    t0 = time.time()
    for it in range(BURNIN + DRAWS):
        R.ref_sweep(*args)
        if it >= BURNIN:
            b = sc[:, col].astype(np.float64)
            prb = sp.ndtr(b[0])
            k = (it - BURNIN) // per
            bsum[k, :P] += b
            bsum[k, P] += prb
        if it % 1000 == 0:
            print(it, f"{time.time() - t0:.0f}s", flush=True)
    bm = bsum / per
    mean = bm.mean(0)
    mcse = bm.std(0, ddof=1) / np.sqrt(BATCHES)
    np.savez_compressed(os.path.join(HERE, "ref_chain_c1.npz"), mean=mean.astype(np.float32),
                        mcse=mcse.astype(np.float32), burnin=BURNIN, draws=DRAWS, batches=BATCHES,
                        df_offset=P, trunc_eps=0.0)


if __name__ == "__main__":
    main()
