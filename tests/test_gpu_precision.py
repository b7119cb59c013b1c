"""Precision form (N = 1,000, P = 8): error of each FP32 / mixed-precision
update against FP64, stated as numbers and bounded.

X'Z, coefficients and CAR precision are measured at full size (grid
91x109x91 shell mask, V = 230,906): X'Z from the BF16 hi+lo MFMA kernel and
from the FP32 FMA kernel, both on the same GPU draws, against an FP64 product
of those draws; the coefficient update (FP32 Cholesky, FP64 refactor below
the pivot threshold) and the FP64 pair sums / Wishart draw against the FP64
oracle fed the same inputs. The latent draws themselves are compared with
the FP64 oracle on the same N and P over the small ellipsoid mask (the error
is per draw, so mask size does not change it, and the FP64 oracle finishes
in seconds there). The printed JSON line is the report.
"""
import json

import numpy as np
import pytest

import oracle
from bsglmm_amd import Layout, Sampler, synth

pytestmark = pytest.mark.gpu
SEED = 0x0DDBA11CAFEF00D5


def spat_coef_from_ranks(beta, lay):
    hi, hs = lay["host_idx"], lay["host_idx_sc"]
    ids = np.flatnonzero(hi >= 0)
    sc = np.zeros((beta.shape[0], lay["totvoxp"]), np.float32)
    sc[:, hs[ids]] = beta[:, hi[ids]]
    return sc


def rel(a, b, floor=1.0):
    e = np.abs(np.asarray(a, np.float64) - b) / np.maximum(floor, np.abs(b))
    return float(e.max()), float(np.sqrt(np.mean(e * e)))


def test_precision_form_errors_against_fp64():
    print("case c3 precision report (N=1000, P=8), full-size X'Z / beta / Lambda", flush=True)
    d = synth.generate("c3")
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    L = Layout(d["mask"], nx, ny, nz)
    P, N = d["covar"].shape
    assert (P, N) == (8, 1000)
    rng = np.random.default_rng(17)
    sc0 = spat_coef_from_ranks(d["beta_true"] + 0.05 * rng.standard_normal(d["beta_true"].shape)
                               .astype(np.float32), lay)
    prec0 = (np.eye(P) * 3.0 + 0.25).astype(np.float32)
    sweep = 9
    S = {}
    for mode in (1, 0):  # MFMA BF16 hi+lo, FP32 FMA
        S[mode] = Sampler(L, d["covar"], d["data"], seed=SEED, xtz_mode=mode)
        S[mode].set_spat_coef(sc0)
        S[mode].set_prec(prec0)
        S[mode].set_sweep_index(sweep)
        S[mode].update_latent()
        S[mode].compute_xtz()
    z = S[1].get_z()
    assert np.array_equal(z, S[0].get_z())  # same counters, same draws
    X = d["covar"].astype(np.float64)
    m64 = X @ z.astype(np.float64)
    scale = np.abs(X) @ np.abs(z.astype(np.float64))
    report = {"N": N, "P": P, "V": L.totvox}
    for mode, name in ((1, "xtz_mfma_bf16x2"), (0, "xtz_fp32")):
        e = np.abs(S[mode].get_xtz() - m64) / scale
        report[name] = {"max_err_over_sum_abs": float(e.max()), "rms": float(np.sqrt(np.mean(e * e)))}
    # the BF16 hi + lo MFMA product fused into the draws (xtz_mode 3, Z never
    # stored): its own draws are those of the stored form up to the FP32
    # rounding of the means (matrix-core means there, FMA here)
    F = Sampler(L, d["covar"], d["data"], seed=SEED, xtz_mode=3, z_store=2)
    F.set_spat_coef(sc0)
    F.set_prec(prec0)
    F.set_sweep_index(sweep)
    F.update_latent()
    e = np.abs(F.get_xtz() - m64) / scale
    report["xtz_mfma_bf16x2_fused"] = {"max_err_over_sum_abs": float(e.max()), "rms": float(np.sqrt(np.mean(e * e)))}
    F.close()
    del z, m64, scale

    # coefficients from the MFMA X'Z, against the FP64 oracle on the same M
    O = oracle.Oracle(lay, d["covar"], d["data"], seed=SEED)
    m_gpu = S[1].get_xtz().astype(np.float64)
    before = sc0.astype(np.float64)
    for colour in (0, 1):
        S[1].update_coef(colour)
    b_gpu = S[1].get_spat_coef()
    ref = O.update_coef(0, m_gpu, prec0, before, sweep)
    ref = O.update_coef(1, m_gpu, prec0, ref, sweep)
    report["beta"] = dict(zip(("max_rel", "rms_rel"), rel(b_gpu, ref)))
    report["beta"]["fp64_refactors"] = S[1].fallback_count()
    S[1].update_car_prec()
    t_gpu = S[1].get_pair_sums()
    t_orc = O.pair_sums(b_gpu)
    report["pair_sums_max_rel"] = float(np.max(np.abs(t_gpu - t_orc)) / np.abs(t_orc).max())
    lam_orc = O.wishart(t_gpu, lay["totvox"] + 19, sweep)
    report["lambda"] = dict(zip(("max_rel", "rms_rel"), rel(S[1].get_prec(), lam_orc, floor=np.abs(lam_orc).max())))

    # latent draws: same N and P, small mask
    ds = synth.generate("c1", nsubs=N, ncovar=P)
    lays = oracle.layout(ds["mask"], *ds["grid"])
    Ls = Layout(ds["mask"], *ds["grid"])
    Ss = Sampler(Ls, ds["covar"], ds["data"], seed=SEED, xtz_mode=1)
    scs = spat_coef_from_ranks(ds["beta_true"], lays)
    Ss.set_spat_coef(scs)
    Ss.set_sweep_index(sweep)
    Ss.update_latent()
    zo = oracle.Oracle(lays, ds["covar"], ds["data"], seed=SEED).update_z(scs, sweep)
    report["z"] = dict(zip(("max_rel", "rms_rel"), rel(Ss.get_z(), zo)))
    print("PRECISION " + json.dumps(report), flush=True)

    assert report["xtz_mfma_bf16x2"]["max_err_over_sum_abs"] < 2e-5
    assert report["xtz_mfma_bf16x2_fused"]["max_err_over_sum_abs"] < 2e-5
    assert report["xtz_fp32"]["max_err_over_sum_abs"] < 1e-5
    assert report["beta"]["max_rel"] < 1e-5
    # FP64 sums of 675,036 pair terms in a different order: n u = 7.5e-11
    assert report["pair_sums_max_rel"] < 1e-10
    assert report["lambda"]["max_rel"] < 1e-6
    assert report["z"]["max_rel"] < 1e-5
