"""Wide problems: 9 <= P <= 20 covariates, the reference's upper range
(per-thread var[20*20], covarGPU.cu:65, 300-304), on the wide kernels
(k_latent_w, k_coef_w, k_pairs_w, k_wishart_w and the per-save summaries),
each update against the FP64 oracle with the same Philox counters.

Tolerances: X'Z within 2e-5 of sum_i |x_i||z_i| (FP32 draws summed in FP32);
coefficients within 1e-5 relative (floor 1) of FP64, as for P <= 8 (measured:
2.5e-6 at P = 10 to 3.5e-6 at P = 20); pair sums within 1e-12; Lambda within
1e-6 of its largest entry; the FP64 refactor within 1e-6 (measured 9e-8).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import bsglmm_amd as B
from bsglmm_amd import Batch, BsgError, Layout, Sampler, synth

pytestmark = pytest.mark.gpu
SEED = 0x9ABCDEF012345678
TOL_COEF = 1e-5


def spat_coef_from_ranks(beta, lay):
    hi, hs = lay["host_idx"], lay["host_idx_sc"]
    ids = np.flatnonzero(hi >= 0)
    sc = np.zeros((beta.shape[0], lay["totvoxp"]), np.float32)
    sc[:, hs[ids]] = beta[:, hi[ids]]
    return sc


def phi_as(x):
    """Abramowitz & Stegun 26.2.17, as the reference's ProbSDNorm (covar.cpp:327-348)."""
    t = 1.0 / (1.0 + 0.2316419 * np.abs(x))
    poly = t * (0.319381530 + t * (-0.356563782 + t * (1.781477937 + t * (-1.821255978 + t * 1.330274429))))
    tail = 0.39894228 * np.exp(-0.5 * x * x) * poly
    return np.where(x >= 0, 1.0 - tail, tail)


def rel_err(a, b, floor=1.0):
    return float(np.max(np.abs(a - b) / np.maximum(floor, np.abs(b))))


def setup(nsubs, ncovar, nchains=1, wm=False, **kw):
    d = synth.generate("c1", nsubs=nsubs, ncovar=ncovar)
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    L = Layout(d["mask"], nx, ny, nz)
    w = np.random.default_rng(4).uniform(0.2, 1.0, L.totvox).astype(np.float32) if wm else None
    S = Sampler(L, d["covar"], d["data"], wm=w, seed=SEED, nchains=nchains, xtz_mode=2, z_store=2, **kw)
    O = oracle.Oracle(lay, d["covar"], d["data"], wm=w, seed=SEED)
    return d, lay, L, S, O, w


def start(S, d, lay, nchains, sweep, rng):
    P = S.P
    prec0 = (np.eye(P) * 1.5 + 0.1).astype(np.float32)
    S.set_sweep_index(sweep)
    scs = []
    for c in range(nchains):
        beta = d["beta_true"] + 0.1 * rng.standard_normal(d["beta_true"].shape).astype(np.float32)
        scs.append(spat_coef_from_ranks(beta, lay))
        S.set_spat_coef(scs[c], chain=c)
        S.set_prec(prec0, chain=c)
    return prec0, scs


@pytest.mark.parametrize("nsubs,ncovar,nchains", [(60, 12, 1), (150, 20, 2), (37, 10, 1), (300, 16, 1)])
def test_wide_sweep_every_update_matches_oracle(nsubs, ncovar, nchains):
    print(f"case wide sweep c1 N={nsubs} P={ncovar} chains={nchains}", flush=True)
    d, lay, L, S, O, _ = setup(nsubs, ncovar, nchains)
    assert not S.overlap
    sweep = 3
    prec0, scs = start(S, d, lay, nchains, sweep, np.random.default_rng(17))
    S.sweep(1)
    S.synchronize()
    assert S.fallback_count() == 0
    ec = 0.0
    for c in range(nchains):
        O.set_chain(c)
        z = O.update_z(scs[c], sweep)
        m = S.get_xtz(c)
        scale = np.abs(O.covar).astype(np.float64) @ np.abs(z)
        assert np.max(np.abs(m - O.xtz(z.astype(np.float32))) / scale) < 2e-5
        ref = O.update_coef(0, m.astype(np.float64), prec0, scs[c].astype(np.float64), sweep)
        ref = O.update_coef(1, m.astype(np.float64), prec0, ref, sweep)
        b = S.get_spat_coef(c)
        e = rel_err(b, ref)
        ec = max(ec, e)
        assert e < TOL_COEF, e
        t = S.get_pair_sums(c)
        to = O.pair_sums(b)
        assert np.max(np.abs(t - to)) <= 1e-12 * np.abs(to).max()
        lam = O.wishart(t, lay["totvox"] + 19, sweep)
        assert rel_err(S.get_prec(c), lam, floor=np.abs(lam).max()) < 1e-6
    print(f"P={ncovar}: coefficients within {ec:.2e} of FP64", flush=True)


def test_wide_stepwise_api_and_summaries():
    """bsg_update_latent / bsg_update_coef / bsg_update_car_prec step by step at
    P = 14, then the per-save summaries: plane means, log-likelihood and the
    posterior accumulators against FP64 evaluations of the GPU's own draws."""
    print("case wide steps + summaries c1 N=80 P=14", flush=True)
    d, lay, L, S, O, _ = setup(80, 14)
    P = 14
    sweep = 9
    prec0, scs = start(S, d, lay, 1, sweep, np.random.default_rng(5))
    S.update_latent()
    z = O.update_z(scs[0], sweep)
    m = S.get_xtz()
    scale = np.abs(O.covar).astype(np.float64) @ np.abs(z)
    assert np.max(np.abs(m - O.xtz(z.astype(np.float32))) / scale) < 2e-5
    with pytest.raises(BsgError):
        S.compute_xtz()  # Z is not kept
    ref = scs[0].astype(np.float64)
    for colour in (0, 1):
        before = S.get_spat_coef()
        S.update_coef(colour)
        ref = O.update_coef(colour, m.astype(np.float64), prec0, before.astype(np.float64), sweep)
        assert rel_err(S.get_spat_coef(), ref) < 1e-5
    S.update_car_prec()
    b = S.get_spat_coef()
    t = S.get_pair_sums()
    assert np.max(np.abs(t - O.pair_sums(b))) <= 1e-12 * np.abs(t).max()
    lam = O.wishart(t, lay["totvox"] + 19, sweep)
    assert rel_err(S.get_prec(), lam, floor=np.abs(lam).max()) < 1e-6
    assert S.sweep_index() == sweep + 1
    # plane means and log-likelihood
    e = L.export()
    hs = e["host_idx_sc"][e["host_idx"] >= 0]
    order = np.argsort(e["host_idx"][e["host_idx"] >= 0])
    bm = b[:, hs].astype(np.float64)
    assert np.allclose(S.coef_means(), bm.mean(1), rtol=1e-10, atol=1e-12)
    T = 3
    ll = S.loglik(ntypes=T)
    bv = bm[:, order]  # mask-rank order
    X = O.covar.astype(np.float64)
    y = d["data"].astype(np.float64)
    rest = X[T:].T @ bv[T:]  # N x V
    for tt in range(T):
        p = phi_as(bv[tt][None, :] + rest)
        ref_ll = np.where(y > 0, np.log(p + 1e-35), np.log(1 - p + 1e-35)).sum(1)
        assert np.allclose(ll[:, tt], ref_ll, rtol=2e-5)
    # accumulators against host snapshots
    S.accum_reset(ntypes=2, nbatches=3, batch_len=2)
    snaps = []
    for _ in range(6):
        S.sweep(1)
        S.accum_add()
        snaps.append(S.get_spat_coef()[:, hs][:, order].astype(np.float64))
    out = S.accum_get()
    sn = np.array(snaps)
    assert out["ndraws"] == 6
    assert np.allclose(out["mean"], sn.mean(0), rtol=1e-6, atol=1e-6)
    assert np.allclose(out["var"], sn.var(0, ddof=1), rtol=1e-4, atol=1e-7)
    # the fused accumulation on save sweeps equals the separate pass
    d2, lay2, L2, F, _, _ = setup(80, 14)
    A = Sampler(L2, d2["covar"], d2["data"], seed=SEED, xtz_mode=2, z_store=2)
    for X_ in (F, A):
        X_.sweep(2)
        X_.accum_reset(ntypes=1, nbatches=2, batch_len=2)
    F.sweep_accumulate(4, 1)
    for _ in range(4):
        A.sweep(1)
        A.accum_add()
    fa, aa = F.accum_get(), A.accum_get()
    for k in ("mean", "var", "prb", "mcse"):
        assert np.array_equal(fa[k], aa[k]), k


@pytest.mark.parametrize("ncovar", [12, 20])
def test_wide_fp64_refactor_and_non_spd(ncovar):
    print(f"case wide FP64 refactor and non-SPD c1 N=120 P={ncovar}", flush=True)
    d, lay, L, S, O, _ = setup(120, ncovar, pivot_min=1e6)
    sweep = 2
    prec0, scs = start(S, d, lay, 1, sweep, np.random.default_rng(8))
    S.sweep(1)
    S.synchronize()
    assert S.fallback_count() == L.totvox
    m = S.get_xtz().astype(np.float64)
    ref = O.update_coef(0, m, prec0, scs[0].astype(np.float64), sweep)
    ref = O.update_coef(1, m, prec0, ref, sweep)
    e = rel_err(S.get_spat_coef(), ref)
    print(f"P={ncovar}: FP64 refactor within {e:.2e} of the oracle", flush=True)
    assert e < 1e-6
    # a precision with a negative eigenvalue: those voxels keep their coefficients
    d, lay, L, S, O, _ = setup(120, ncovar)
    sc0 = spat_coef_from_ranks(d["beta_true"], lay)
    S.set_spat_coef(sc0)
    S.set_prec((-1000.0 * np.eye(ncovar)).astype(np.float32))
    S.sweep(1)
    with pytest.raises(BsgError, match="positive definite"):
        S.synchronize()
    S.synchronize()
    after = S.get_spat_coef()
    e = L.export()
    hs = e["host_idx_sc"][e["host_idx"] >= 0]
    kept = np.all(after[:, hs] == sc0[:, hs], axis=0)
    assert kept.mean() > 0.9


def test_wide_wm_coefficient():
    print("case wide WM coefficient c1 N=60 P=10", flush=True)
    d, lay, L, S, O, w = setup(60, 10, wm=True)
    sweep, b0 = 7, 0.2
    sc0 = spat_coef_from_ranks(d["beta_true"], lay)
    prec0 = (np.eye(10) * 1.5 + 0.1).astype(np.float32)
    S.set_spat_coef(sc0)
    S.set_prec(prec0)
    S.set_wm_coef(b0)
    S.set_sweep_index(sweep)
    S.sweep(1)
    S.synchronize()
    z = O.update_z(sc0, sweep, b=b0)
    s_orc, w_orc = O.wm_stat(sc0, z)
    b_orc = O.wm_draw(s_orc, w_orc, sweep)
    b_gpu = S.get_wm_coef()
    assert abs(b_gpu - b_orc) < 1e-6 + 1e-5 * abs(b_orc), (b_gpu, b_orc)
    m_ref = O.xtz(z.astype(np.float32), b_gpu)
    ref = O.update_coef(0, m_ref, prec0, sc0.astype(np.float64), sweep)
    ref = O.update_coef(1, m_ref, prec0, ref, sweep)
    assert rel_err(S.get_spat_coef(), ref) < 2e-5


def test_wide_graph_checkpoint_and_full_size(tmp_path):
    """HIP-graph replays equal direct sweeps; a checkpoint resumes exactly; the
    full-brain C2 mask at P = 20 stays finite with a positive definite Lambda
    and repeats bit for bit."""
    print("case wide graph + checkpoint c1 N=50 P=18 x2 chains; c2 mask N=250 P=20", flush=True)
    d = synth.generate("c1", nsubs=50, ncovar=18)
    L = Layout(d["mask"], *d["grid"])
    A, G = (Sampler(L, d["covar"], d["data"], seed=SEED, nchains=2, xtz_mode=2, z_store=2) for _ in range(2))
    for step in (10, 3, 8):
        A.sweep(step)
        if step == 3:
            G.sweep(step)
        else:
            G.sweep_graph(step, 4)
        A.synchronize()
        G.synchronize()
        for c in range(2):
            assert np.array_equal(A.get_spat_coef(c), G.get_spat_coef(c))
            assert np.array_equal(A.get_prec(c), G.get_prec(c))
    path = os.path.join(tmp_path, "wide.bsg")
    A.checkpoint(path)
    R = Sampler(L, d["covar"], d["data"], seed=1, nchains=2, xtz_mode=2, z_store=2)
    R.restore(path)
    A.sweep(5)
    R.sweep(5)
    for c in range(2):
        assert np.array_equal(A.get_spat_coef(c), R.get_spat_coef(c))
    d = synth.generate("c2", ncovar=20)
    L = Layout(d["mask"], *d["grid"])
    runs = []
    for _ in range(2):
        S = Sampler(L, d["covar"], d["data"], seed=SEED, xtz_mode=2, z_store=2)
        S.sweep(3)
        S.synchronize()
        runs.append((S.get_spat_coef(), S.get_prec()))
        S.close()
    (b0, l0), (b1, l1) = runs
    assert np.array_equal(b0, b1) and np.array_equal(l0, l1)
    assert np.all(np.isfinite(b0)) and np.all(np.linalg.eigvalsh(l0.astype(np.float64)) > 0)


def test_wide_limits():
    d = synth.generate("c1", nsubs=20, ncovar=12)
    L = Layout(d["mask"], *d["grid"])
    with pytest.raises(BsgError, match="Z-free"):
        Sampler(L, d["covar"], d["data"], seed=1, xtz_mode=0, z_store=0)
    S = Sampler(L, d["covar"], d["data"], seed=1, xtz_mode=2, z_store=2)
    with pytest.raises(BsgError, match="at most"):
        Batch([S])
    d21 = synth.generate("c1", nsubs=20, ncovar=21)
    with pytest.raises(BsgError, match="1..20"):
        Sampler(L, d21["covar"], d21["data"], seed=1, xtz_mode=2, z_store=2)


def test_bincar_with_twenty_covariates(tmp_path):
    """BinCar with NCov = 20 (two groups and 18 further covariates) runs to the
    end and writes every map (the reference accepts NCov <= 20)."""
    print("case BinCar c1 N=40 NTypes=2 NCov=20, 1200/200", flush=True)
    from nifti_fixture import make_nifti
    d = synth.generate("c1", nsubs=40, ncovar=19)
    nx, ny, nz = d["grid"]
    mask = d["mask"].reshape(nz + 2, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1].astype(np.float64)
    img = tmp_path / "images"
    img.mkdir()
    make_nifti(str(img / "mask.nii.gz"), mask, 64)
    L = Layout(d["mask"], nx, ny, nz)
    e = L.export()
    inner = e["host_idx"].reshape(nz + 2, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1]
    inside = inner >= 0
    grp = (np.arange(40) >= 20).astype(np.float32)
    covar = np.vstack([1 - grp, grp, d["covar"][1:]]).astype(np.float32)
    names = ["gA", "gB"] + [f"x{k}" for k in range(18)]
    rows = ["ID " + " ".join(names) + " Image"]
    for i in range(40):
        vol = np.zeros(mask.shape, np.float64)
        vol[inside] = d["data"][i][inner[inside]]
        make_nifti(str(img / f"s{i}.nii.gz"), vol, 16)
        rows.append(f"sub{i} " + " ".join(repr(float(v)) for v in covar[:, i]) + f" s{i}.nii.gz")
    (tmp_path / "design.txt").write_text("\n".join(rows) + "\n")
    binc = os.path.join(os.path.dirname(B.LIB_PATH), "BinCar")
    r = subprocess.run([binc, "2", "20", "1", "design.txt", "1", "0", "1200", "200"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-600:], flush=True)
    assert r.returncode == 0, r.stdout[-1500:]
    for nm in names:
        assert (tmp_path / f"spatCoef_{nm}.nii.gz").exists(), nm
    assert (tmp_path / "DIC.dat").exists() and (tmp_path / "Qhat.dat").exists()
