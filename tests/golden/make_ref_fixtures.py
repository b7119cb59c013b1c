"""Writes golden vectors from the reference's own CPU routines (oracle/_ref,
i.e. bsglmm/covar.cpp, randgen.cpp, cholesky.cpp compiled unchanged), plus
the bundled mask as a bit-packed array. Run once where the reference tree
and oracle/_ref exist:

    make -C oracle && python tests/golden/make_ref_fixtures.py

Fixtures (all small):
  ref_kiss.json        KISS stream and devseed from the seed.dat triple (randgen.cpp:9-45, main.cu:413)
  ref_truncnorm.npz    empirical CDFs of 200,000 truncNorm2 draws per (mu, y) setting (randgen.cpp:726-814)
  ref_update.npz       one updateSpatCoef + updateSpatCoefPrec on a small problem, with the
                       normal / chi-square variates the reference consumed (covar.cpp:121-299,
                       cholesky.cpp:245-286)
  bundled_mask.npz     bsglmm/images/mask.nii.gz as a padded, bit-packed byte grid
"""
import ctypes as C
import gzip
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

REF_TREE = "/root/reference/bsglmm"
SEED_DAT = (5118772087888418775, 289433066200759127, 1964621368)  # bsglmm/seed.dat
TN_SETTINGS = [(-2.0, 0), (-2.0, 1), (0.0, 0), (0.0, 1), (0.3, 1), (1.5, 1), (-4.0, 1), (2.0, 0)]


def small_problem():
    """10x10x10 grid, ellipsoid mask, N=30, P=3 (also rebuilt by the tests)."""
    nx = ny = nz = 10
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    m = ((x - 4.5) / 4.5) ** 2 + ((y - 4.5) / 4.0) ** 2 + ((z - 4.5) / 3.5) ** 2 <= 1.0
    return np.pad(m.astype(np.uint8), 1), (nx, ny, nz)


def kiss_fixture(R):
    s = oracle.seed_array(SEED_DAT)
    draws = [R.ref_kiss(s) for _ in range(1000)]
    s2 = oracle.seed_array(SEED_DAT)
    dev = R.ref_runiform_long_n(C.c_ulong(2 ** 64 - 1), s2)
    return {"seed": list(SEED_DAT), "draws_hex": [float(d).hex() for d in draws],
            "state_after": [int(v) for v in s], "devseed": int(dev), "state_after_devseed": [int(v) for v in s2]}


def truncnorm_fixture(R, n=200000):
    s = oracle.seed_array(SEED_DAT)
    grids, cdfs = [], []
    for mu, y in TN_SETTINGS:
        lo, hi = (0.0, 500.0) if y else (-500.0, -0.0)
        xs = np.array([R.ref_truncnorm2(mu, 1.0, lo, hi, s) for _ in range(n)])
        q = np.quantile(xs, np.linspace(0.001, 0.999, 200))
        grids.append(q)
        cdfs.append(np.searchsorted(np.sort(xs), q, side="right") / n)
    return dict(settings=np.array(TN_SETTINGS), n=n, grid=np.array(grids), ecdf=np.array(cdfs))


def update_fixture(R):
    mask, (nx, ny, nz) = small_problem()
    lay = oracle.layout(mask, nx, ny, nz)
    V, VP = lay["totvox"], lay["totvoxp"]
    rng = np.random.default_rng(1203)
    N, P = 30, 3
    covar = np.array([np.ones(N), rng.standard_normal(N), (rng.random(N) < 0.5)], np.float32)
    Z = (rng.standard_normal((N, V)) - 1.0).astype(np.float32)
    sc = np.zeros((P, VP), np.float32)
    hi, hs = lay["host_idx"], lay["host_idx_sc"]
    ids = np.flatnonzero(hi >= 0)
    sc[:, hs[ids]] = (0.3 * rng.standard_normal((P, len(ids))) - np.array([[1.0], [0.0], [0.0]])).astype(np.float32)
    A = rng.standard_normal((P, P))
    prec = (A @ A.T + P * np.eye(P)).astype(np.float32)
    xtx = oracle.xtx_reference(covar)
    oracle.RefCPU(lay, covar, xtx)
    wm = np.ones(V, np.float32)
    f = C.POINTER(C.c_float)
    # the normals updateSpatCoef will consume: P per voxel, colour-0 list then colour 1
    s_norm = oracle.seed_array(SEED_DAT)
    normals = np.array([R.ref_snorm(s_norm) for _ in range(V * P)])
    sc_after = sc.copy()
    s = oracle.seed_array(SEED_DAT)
    R.ref_update_spat_coef(covar.ctypes.data_as(f), sc_after.ctypes.data_as(f), prec.ctypes.data_as(f),
                           Z.ctypes.data_as(f), 0.0, wm.ctypes.data_as(f), s)
    assert list(s) == list(s_norm), "updateSpatCoef consumed a different number of draws"
    # pair sums per (a >= b), both colours (covar.cpp:276-283 minus the +I)
    T = np.zeros((P, P))
    for a in range(P):
        for b in range(a + 1):
            acc = 0.0
            for c in range(2):
                acc += R.ref_spat_coef_precision(sc_after[a].ctypes.data_as(f), sc_after[b].ctypes.data_as(f), c)
            T[a, b] = T[b, a] = acc
    df = V + P  # covar.cpp:291
    s_var = oracle.seed_array(list(s))
    chi2, norms = np.zeros(P), np.zeros((P, P))
    for i in range(P):  # rwishart2's order of draws (cholesky.cpp:259-263)
        chi2[i] = R.ref_rgamma(0.5 * (df - i), 0.5, s_var)
        for j in range(i):
            norms[i, j] = R.ref_snorm(s_var)
    lam = prec.copy()
    R.ref_update_spat_coef_prec(lam.ctypes.data_as(f), sc_after.ctypes.data_as(f), s)
    assert list(s) == list(s_var), "updateSpatCoefPrec consumed a different number of draws"
    return dict(N=N, P=P, covar=covar, Z=Z, sc_before=sc, prec=prec, normals=normals, sc_after=sc_after,
                T=T, df=df, chi2=chi2, wnorms=norms, lam=lam)


def read_nifti_f64(path):
    with gzip.open(path, "rb") as fh:
        raw = fh.read()
    hdr = np.frombuffer(raw[:348], dtype=np.uint8)
    dim = np.frombuffer(raw[40:56], dtype="<i2")
    dtype = int(np.frombuffer(raw[70:72], dtype="<i2")[0])
    off = int(np.frombuffer(raw[108:112], dtype="<f4")[0])
    assert dtype == 64, "mask must be FLOAT64 (read_data.cpp:204)"
    nx, ny, nz = int(dim[1]), int(dim[2]), int(dim[3])
    vals = np.frombuffer(raw[off:off + 8 * nx * ny * nz], dtype="<f8").reshape(nz, ny, nx)
    del hdr
    return vals, (nx, ny, nz)


def mask_fixture():
    vals, (nx, ny, nz) = read_nifti_f64(os.path.join(REF_TREE, "images", "mask.nii.gz"))
    padded = np.pad(vals.astype(np.uint8), 1)  # (unsigned char) of the double, read_data.cpp:235
    return dict(grid=np.array([nx, ny, nz]), bits=np.packbits(padded.reshape(-1) != 0),
                nsize=padded.size)


def main():
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle)")
    with open(os.path.join(HERE, "ref_kiss.json"), "w") as fh:
        json.dump(kiss_fixture(R), fh)
    np.savez_compressed(os.path.join(HERE, "ref_truncnorm.npz"), **truncnorm_fixture(R))
    np.savez_compressed(os.path.join(HERE, "ref_update.npz"), **update_fixture(R))
    np.savez_compressed(os.path.join(HERE, "bundled_mask.npz"), **mask_fixture())


if __name__ == "__main__":
    main()
