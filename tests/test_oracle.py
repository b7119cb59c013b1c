"""Pins the FP64 oracle (CPU only).

Checks are chosen so that a dropped term, a wrong sign, index or truncation
side, or a transposed factor fails one of them:
  * Philox4x32-10 against the Random123 known-answer vectors and rocRAND;
  * the normal tail and its inverse against scipy (an independent library);
  * the latent draw against scipy's truncated-normal quantile function, and
    its distribution against the reference's own truncNorm2 (fixture);
  * the coefficient update against the reference's updateSpatCoef given the
    normals it consumed (fixture), and against a closed form (numpy solve and
    the Cholesky factor of the inverse) on random problems;
  * the pair sums and the Wishart draw against the reference's
    updateSpatCoefPrec given the variates it consumed (fixture), and the
    Wishart and gamma samplers by their moments;
  * KISS against the reference's stream (fixture).
"""
import ctypes as C
import json
import os

import numpy as np
import pytest
import scipy.special as sp
import scipy.stats as st

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def u01(x):
    return ((np.asarray(x, np.uint64) >> 9) * 2 + 1) * 2.0 ** -24


# ---------------------------------------------------------------- Philox
def test_philox_random123_known_answers():
    # Salmon et al. (SC'11), Random123 kat_vectors for philox4x32_10
    kat = [([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
           ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
           ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
            [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1])]
    for ctr, key, out in kat:
        assert list(oracle.philox(ctr, key)) == out


def test_philox_matches_rocrand():
    with open(os.path.join(GOLD, "philox_rocrand.json")) as fh:
        cases = json.load(fh)["cases"]
    for c in cases:
        assert list(oracle.philox(c["counter"], c["key"])) == c["out"]


def test_counter_layout_and_uniform_grid():
    key = np.array([7, 9], np.uint32)
    out = np.empty(4, np.uint32)
    o = oracle.orc()
    o.orc_draw(key.ctypes.data_as(C.POINTER(C.c_uint32)), 123, 45, 1, 6, 3,
               out.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert list(out) == list(oracle.philox([123, 45, 1 | (6 << 4), 3], [7, 9]))
    xs = np.array([0, 1, 511, 512, 2 ** 31, 2 ** 32 - 1], np.uint32)
    for x in xs:
        u = o.orc_u01(int(x))
        assert 0 < u < 1 and u == np.float64(np.float32(u))  # exact in float32
        assert u == u01(x)
    assert o.orc_u01(0) == 2.0 ** -24 and o.orc_u01(2 ** 32 - 1) == 1 - 2.0 ** -24


# ---------------------------------------------------------- normal tail
def test_normal_tail_against_scipy():
    o = oracle.orc()
    for x in np.concatenate([np.linspace(-8, 8, 161), [10, 20, 29.9, 30.1, 35, 40]]):
        assert abs(o.orc_logQ(float(x)) - sp.log_ndtr(-x)) <= 1e-13 * max(1, abs(sp.log_ndtr(-x)))
    ps = np.concatenate([10.0 ** -np.arange(300, 1, -7), np.linspace(0.01, 0.99, 99)])
    for p in ps:
        ref = -sp.ndtri(p)
        assert abs(o.orc_Qinv(float(p)) - ref) <= 1e-13 * max(1, abs(ref))


def test_latent_draw_is_the_truncated_normal_quantile():
    o = oracle.orc()
    for mu in (-6, -2.5, -2, -0.5, 0.0, 0.3, 2, 5, 9, 20):
        for y in (0, 1):
            for eps in (1e-6, 0.0):
                lo, hi = (eps, 500.0) if y else (-500.0, -eps)
                for u in (u01(1), 0.01, 0.3, 0.5, 0.77, 0.999, u01(2 ** 32 - 1)):
                    z = o.orc_latent_draw(mu, y, u, eps)
                    a, b = lo - mu, hi - mu
                    # u is the upper-tail probability of W = s (Z - mu)
                    ref = st.truncnorm.isf(u, a, b, loc=mu) if y else st.truncnorm.ppf(u, a, b, loc=mu)
                    assert abs(z - ref) <= 1e-8 * max(1, abs(ref)), (mu, y, u, z, ref)
                    assert (z > lo) if y else (z < hi)


def test_latent_draws_match_reference_truncnorm_distribution():
    """Empirical CDFs of 200,000 draws of the reference's truncNorm2 (Robert 1995,
    limits (0,500) / (-500,-0) as covar.cpp:317,320) against the exact truncated
    normal CDF and against 200,000 oracle inversion draws driven by Philox."""
    f = np.load(os.path.join(GOLD, "ref_truncnorm.npz"))
    n = int(f["n"])
    o = oracle.orc()
    key = np.array([0x1234, 0x5678], np.uint32)
    kp = key.ctypes.data_as(C.POINTER(C.c_uint32))
    out = np.empty(4, np.uint32)
    op = out.ctypes.data_as(C.POINTER(C.c_uint32))
    crit = 1.95 * np.sqrt(2.0 / n)  # two-sample KS, alpha = 1e-3
    for k, ((mu, y), grid, ecdf) in enumerate(zip(f["settings"], f["grid"], f["ecdf"])):
        y = int(y)
        lo, hi = (0.0, 500.0) if y else (-500.0, 0.0)
        if mu == 0.0 and y == 0:
            # Reference quirk (randgen.cpp:737-774): with limits (-500, -0.0) and
            # mu = 0 the standardised upper limit is -0.0, which takes the
            # unswapped branch (v < 0 is false) and is then negated (v <= 0 is
            # true), so every draw is positive. The CPU path meets it in its first
            # sweep (beta = 0); the GPU limits +-1e-6 (covarGPU.cu:1414-1418) avoid
            # it, and the oracle follows the stated truncation.
            assert np.all(grid > 0)
            continue
        exact = st.truncnorm.cdf(grid, lo - mu, hi - mu, loc=mu)
        assert np.max(np.abs(ecdf - exact)) < 1.63 / np.sqrt(n) + 1e-4, (mu, y)
        draws = np.empty(n)
        for g in range(n // 4):
            o.orc_draw(kp, g, k, 1, 0, 0, op)
            for j in range(4):
                draws[4 * g + j] = o.orc_latent_draw(float(mu), y, u01(out[j]), 0.0)
        ours = np.searchsorted(np.sort(draws), grid, side="right") / n
        assert np.max(np.abs(ours - ecdf)) < crit, (mu, y)


# -------------------------------------------------------- coefficient update
def small_problem():
    nx = ny = nz = 10
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    m = ((x - 4.5) / 4.5) ** 2 + ((y - 4.5) / 4.0) ** 2 + ((z - 4.5) / 3.5) ** 2 <= 1.0
    return np.pad(m.astype(np.uint8), 1), (nx, ny, nz)


@pytest.fixture(scope="module")
def ref_update():
    f = dict(np.load(os.path.join(GOLD, "ref_update.npz")))
    mask, (nx, ny, nz) = small_problem()
    lay = oracle.layout(mask, nx, ny, nz)
    V = lay["totvox"]
    data = np.zeros((int(f["N"]), V), np.uint8)
    O = oracle.Oracle(lay, f["covar"], data)
    return f, lay, O


def test_coef_update_matches_reference(ref_update):
    """The reference's updateSpatCoef (FP32: covar.cpp:121-187) against the FP64
    oracle, both given the same Z, Lambda and the normals the reference drew.
    Tolerance 2e-5 relative (floor 1): the reference's own FP32 rounding."""
    f, lay, O = ref_update
    m = O.xtz(f["Z"])
    n0, P = lay["n0"], int(f["P"])
    z = f["normals"].reshape(-1, P)
    sc = O.update_coef(0, m, f["prec"], f["sc_before"].astype(np.float64), 0, z_inj=z[:n0])
    sc = O.update_coef(1, m, f["prec"], sc, 0, z_inj=z[n0:])
    err = np.max(np.abs(sc - f["sc_after"]) / np.maximum(1, np.abs(f["sc_after"])))
    assert err < 2e-5, err
    # a transposed Cholesky factor or a swapped colour order is far outside that
    assert np.max(np.abs(f["sc_after"] - f["sc_before"])) > 0.05


def test_pair_sums_match_reference(ref_update):
    """covar.cpp:220-262 (float products, double sums) against the oracle (exact
    FP64 products of the float differences): 1e-6 relative."""
    f, lay, O = ref_update
    T = O.pair_sums(f["sc_after"])
    assert np.max(np.abs(T - f["T"])) <= 1e-6 * np.abs(f["T"]).max()


def test_wishart_matches_reference(ref_update):
    """updateSpatCoefPrec (covar.cpp:264-299 with rwishart2, cholesky.cpp:245-286)
    against the oracle fed the chi-square and normal variates the reference drew."""
    f, lay, O = ref_update
    T = O.pair_sums(f["sc_after"])
    lam = oracle.wishart(int(f["P"]), T, int(f["df"]), [0, 0], 0, chi2=f["chi2"], norms=f["wnorms"])
    assert np.max(np.abs(lam - f["lam"])) <= 2e-6 * np.abs(f["lam"]).max()


def test_coef_update_closed_form():
    """beta = Q^-1 r + chol_lower(Q^-1) z, with Q and r built independently."""
    rng = np.random.default_rng(3)
    mask, (nx, ny, nz) = small_problem()
    lay = oracle.layout(mask, nx, ny, nz)
    V, VP, P, N = lay["totvox"], lay["totvoxp"], 4, 25
    covar = rng.standard_normal((P, N)).astype(np.float32)
    O = oracle.Oracle(lay, covar, np.zeros((N, V), np.uint8))
    m = rng.standard_normal((P, V))
    A = rng.standard_normal((P, P))
    prec = (A @ A.T + np.eye(P)).astype(np.float32)
    sc = np.zeros((P, VP))
    hi, hs = lay["host_idx"], lay["host_idx_sc"]
    ids = np.flatnonzero(hi >= 0)
    sc[:, hs[ids]] = rng.standard_normal((P, len(ids)))
    z = rng.standard_normal((lay["n0"], P))
    out = O.update_coef(0, m, prec, sc, 0, z_inj=z)
    xtx = O.xtx.astype(np.float64)
    sx, sxy = nx + 2, (nx + 2) * (ny + 2)
    for q in range(0, lay["n0"], 17):
        vox = lay["vox0"][q]
        nbrs = [vox - 1, vox + 1, vox - sx, vox + sx, vox - sxy, vox + sxy]
        sumv = sum(sc[:, hs[k]] for k in nbrs)
        nn = sum(mask.reshape(-1)[k] != 0 for k in nbrs)
        Q = xtx + nn * prec.astype(np.float64)
        r = m[:, hi[vox]] + prec.astype(np.float64) @ sumv
        L = np.linalg.cholesky(np.linalg.inv(Q))
        ref = np.linalg.solve(Q, r) + L @ z[q]
        assert np.allclose(out[:, hs[vox]], ref, rtol=1e-11, atol=1e-11)


# ------------------------------------------------------ Wishart and gamma
def test_gamma_sampler_distribution():
    key = np.array([11, 13], np.uint32)
    kp = key.ctypes.data_as(C.POINTER(C.c_uint32))
    o = oracle.orc()
    for a in (1.0, 3.5, 50.0, 115463.5):
        xs = np.array([o.orc_gamma(a, kp, 0, s, 0) for s in range(20000)])
        assert st.kstest(xs, st.gamma(a).cdf).pvalue > 1e-3, a


def test_wishart_mean_is_df_times_scale():
    rng = np.random.default_rng(5)
    P, df = 3, 40
    A = rng.standard_normal((P, 3 * P))
    T = A @ A.T
    S = np.linalg.inv(T + np.eye(P))
    draws = np.array([oracle.wishart(P, T, df, [1, 2], s) for s in range(20000)])
    mean = draws.mean(0)
    var = df * (S ** 2 + np.outer(np.diag(S), np.diag(S)))  # Var(W_ij) for a Wishart
    assert np.all(np.abs(mean - df * S) < 5 * np.sqrt(var / len(draws)))


# ------------------------------------------------------------------ KISS
def test_kiss_matches_reference_stream():
    with open(os.path.join(GOLD, "ref_kiss.json")) as fh:
        f = json.load(fh)
    o = oracle.orc()
    s = oracle.seed_array(f["seed"])
    for h in f["draws_hex"]:
        assert o.orc_kiss(s) == float.fromhex(h)
    assert list(s) == f["state_after"]
    s = oracle.seed_array(f["seed"])
    assert o.orc_runiform_long_n(C.c_ulong(2 ** 64 - 1), s) == f["devseed"]
    assert list(s) == f["state_after_devseed"]


def test_wm_stat_closed_form_and_draw_law():
    """orc_wm_stat/orc_wm_draw (covarGPU.cu:588-724): with Z = x'b_j + c WM_j
    exactly, S/W = c (the residual is c WM_j, so S = c N sum WM^2 = c W);
    with a flipped sign or a wrong covariate index S/W would miss c. The draws
    over many sweeps follow N(S/W, 1/W)."""
    from bsglmm_amd import synth
    d = synth.generate("c1", nsubs=12, ncovar=3)
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    rng = np.random.default_rng(3)
    V, N = lay["totvox"], 12
    wm = rng.uniform(0.1, 1.0, V).astype(np.float32)
    O = oracle.Oracle(lay, d["covar"], d["data"], wm=wm, seed=99)
    beta = rng.standard_normal((3, V)).astype(np.float32)
    hi, hs = lay["host_idx"], lay["host_idx_sc"]
    ids = np.flatnonzero(hi >= 0)
    sc = np.zeros((3, lay["totvoxp"]), np.float32)
    sc[:, hs[ids]] = beta[:, hi[ids]]
    c = -0.37
    z = (d["covar"].astype(np.float64).T @ beta.astype(np.float64)) + c * wm.astype(np.float64)[None, :]
    S, W = O.wm_stat(sc, z)
    assert W == pytest.approx(N * np.sum(wm.astype(np.float64) ** 2), rel=1e-12)
    assert S / W == pytest.approx(c, rel=1e-9)
    draws = np.array([O.wm_draw(S, W, sw) for sw in range(4000)])
    zs = (draws - S / W) * np.sqrt(W)
    assert st.kstest(zs, "norm").pvalue > 1e-3


def test_cpu_wm_stat_closed_form_and_fp64_agreement():
    """orc_cpu_wm_stat (the reference CPU path's updateBeta statistic,
    mcmc.cpp:1099-1136, float Kahan sums in raster order): with
    Z = x'b_j + c WM_j the mean is c W / (W + prior_prec) and the variance
    1 / (W + prior_prec); with random Z it agrees with the FP64 statistic
    (orc_wm_stat) to float accuracy, and a prior mean shifts the mean by
    var * prior_mean * prior_prec."""
    from bsglmm_amd import synth
    d = synth.generate("c1", nsubs=20, ncovar=3)
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    rng = np.random.default_rng(8)
    V, N = lay["totvox"], 20
    wm = rng.uniform(0.1, 1.0, V).astype(np.float32)
    O = oracle.Oracle(lay, d["covar"], d["data"], wm=wm, seed=5)
    beta = rng.standard_normal((3, V)).astype(np.float32)
    hi, hs = lay["host_idx"], lay["host_idx_sc"]
    ids = np.flatnonzero(hi >= 0)
    sc = np.zeros((3, lay["totvoxp"]), np.float32)
    sc[:, hs[ids]] = beta[:, hi[ids]]
    W = N * np.sum(wm.astype(np.float64) ** 2)
    c = 0.81
    z = ((d["covar"].astype(np.float64).T @ beta.astype(np.float64)) + c * wm[None, :]).astype(np.float32)
    m, v = O.cpu_wm_stat(sc, z)
    assert v == pytest.approx(1.0 / (W + 1e-6), rel=1e-6)
    assert m == pytest.approx(c * W / (W + 1e-6), abs=2e-4)
    zr = rng.standard_normal((N, V)).astype(np.float32)
    m, v = O.cpu_wm_stat(sc, zr)
    S, W2 = O.wm_stat(sc, zr.astype(np.float64))
    assert m == pytest.approx(S / (W2 + 1e-6), abs=1e-4 * np.sqrt(np.sum(np.abs(zr)) / W2))
    m2, _ = O.cpu_wm_stat(sc, zr, prior_mean=3.0, prior_prec=0.5)
    m3, v3 = O.cpu_wm_stat(sc, zr, prior_mean=0.0, prior_prec=0.5)
    assert m2 - m3 == pytest.approx(v3 * 1.5, rel=1e-4)


@pytest.mark.parametrize("mu,y", [(-2.0, 0), (-2.0, 1), (0.3, 0), (0.3, 1), (3.5, 0), (-0.5, 1)])
def test_accept_or_fallback_sampler_is_the_truncated_normal(mu, y):
    """orc_update_z method 1 (Box-Muller candidate kept inside the interval
    of y, exact inversion from an independent block otherwise) against the
    exact truncated-normal CDF: one voxel-major problem with an intercept-only
    design, every subject at the same mean, many sweeps. A sampler that kept
    rejected candidates, clipped them, or reused the candidate's uniform for
    the redraw fails this."""
    nsub = 400
    mask = np.zeros((4, 4, 5), np.uint8)
    mask[1:3, 1:3, 1:4] = 1  # 2x2x3 voxels
    lay = oracle.layout(mask.reshape(-1), 3, 2, 2)
    V = lay["totvox"]
    covar = np.ones((1, nsub), np.float32)
    data = np.full((nsub, V), y, np.uint8)
    O = oracle.Oracle(lay, covar, data, seed=1234)
    sc = np.full((1, lay["totvoxp"]), mu, np.float32)
    z = np.concatenate([O.update_z(sc, sw, method=1).ravel() for sw in range(6)])
    lo, hi = (1e-6, 500.0) if y else (-500.0, -1e-6)
    assert np.all((z > lo) & (z < hi))
    p = st.kstest(z, lambda t: st.truncnorm.cdf(t, lo - mu, hi - mu, loc=mu)).pvalue
    assert p > 1e-3, p
