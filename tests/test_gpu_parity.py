"""GPU parity: each conditional update of one sweep, HIP path vs the FP64 oracle.

Both sides draw their variates from Philox4x32-10 with the same counters, so
the uniforms are bit-identical; each update is fed the GPU's own state from
just before it, and its output is compared with the oracle's FP64 result.
Tolerances (stated per test): 1e-5 relative, with a floor of 1 on the scale
for Z, the coefficients and X'Z (which is compared against the norm
sum_i |x_i||z_i| because it is a sum of N products).
"""
import numpy as np
import pytest

import oracle
from bsglmm_amd import Layout, Sampler, synth

pytestmark = pytest.mark.gpu

SEED = 0x1234ABCD5678EF01


def spat_coef_from_ranks(beta, lay):
    """beta[P][TOTVOX] (mask rank) -> spatCoef[P][TOTVOXp] (dilated rank)."""
    hi, hs = lay["host_idx"], lay["host_idx_sc"]
    ids = np.flatnonzero(hi >= 0)
    sc = np.zeros((beta.shape[0], lay["totvoxp"]), np.float32)
    sc[:, hs[ids]] = beta[:, hi[ids]]
    return sc


def problem(name="c1", nsubs=None, ncovar=None, nchains=1, xtz_mode=0):
    d = synth.generate(name, nsubs=nsubs, ncovar=ncovar)
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    L = Layout(d["mask"], nx, ny, nz)
    S = Sampler(L, d["covar"], d["data"], seed=SEED, nchains=nchains, xtz_mode=xtz_mode)
    O = oracle.Oracle(lay, d["covar"], d["data"], seed=SEED)
    rng = np.random.default_rng(7)
    beta = d["beta_true"] + 0.1 * rng.standard_normal(d["beta_true"].shape).astype(np.float32)
    return d, lay, L, S, O, spat_coef_from_ranks(beta, lay)


def rel_err(a, b, floor=1.0):
    return np.max(np.abs(a - b) / np.maximum(floor, np.abs(b)))


@pytest.mark.parametrize("xtz_mode", [0, 1, 2])
@pytest.mark.parametrize("name,nsubs,ncovar", [("c1", None, None), ("c1", 37, 8), ("c1", 300, 4),
                                               ("c1", 5, 2), ("c1", 200, 8), ("c1", 160, 6),
                                               ("c1", 140, 1)])
def test_one_sweep_every_update_matches_oracle(name, nsubs, ncovar, xtz_mode):
    print(f"case {name} N={nsubs} P={ncovar} xtz_mode={xtz_mode}", flush=True)
    d, lay, L, S, O, sc0 = problem(name, nsubs, ncovar, xtz_mode=xtz_mode)
    P = S.P
    sweep = 11
    S.set_spat_coef(sc0)
    S.set_sweep_index(sweep)
    prec0 = (np.eye(P) * 2.0 + 0.3).astype(np.float32)
    S.set_prec(prec0)

    # latent draws
    S.update_latent()
    z_gpu = S.get_z()
    z_orc = O.update_z(sc0, sweep)
    assert rel_err(z_gpu, z_orc) < 1e-5
    y = d["data"] == 1
    assert np.all(z_gpu[y] > 0) and np.all(z_gpu[~y] < 0)

    # X'Z on the GPU's own Z (mode 2 formed it inside the latent kernel)
    if xtz_mode != 2:
        S.compute_xtz()
    m_gpu = S.get_xtz()
    m_orc = O.xtz(z_gpu)
    scale = np.abs(O.covar).astype(np.float64) @ np.abs(z_gpu.astype(np.float64))
    # FP32 FMA: 1e-5 of sum |x||z|; BF16 hi+lo on MFMA (xtz_mode 1): 2e-5, from
    # the 2^-18 split residual of both operands plus FP32 accumulation
    assert np.max(np.abs(m_gpu - m_orc) / scale) < (2e-5 if xtz_mode == 1 else 1e-5)

    # coefficient updates, colour 0 then colour 1, on the same M
    sc = sc0.astype(np.float64)
    for colour in (0, 1):
        before = S.get_spat_coef()
        S.update_coef(colour)
        after = S.get_spat_coef()
        ref = O.update_coef(colour, m_gpu.astype(np.float64), prec0, before.astype(np.float64), sweep)
        assert rel_err(after, ref) < 1e-5, f"colour {colour}"
        sc = after
    assert S.fallback_count() == 0

    # CAR precision: FP64 pair sums and the Wishart draw
    S.update_car_prec()
    t_gpu = S.get_pair_sums()
    t_orc = O.pair_sums(sc)
    assert np.max(np.abs(t_gpu - t_orc) / np.abs(t_orc).max()) < 1e-12
    lam_gpu = S.get_prec()
    lam_orc = O.wishart(t_gpu, lay["totvox"] + 19, sweep)
    assert rel_err(lam_gpu, lam_orc, floor=np.abs(lam_orc).max()) < 1e-6
    assert S.sweep_index() == sweep + 1
    S.synchronize()


@pytest.mark.parametrize("xtz_mode", [0, 1, 2])
@pytest.mark.parametrize("nsubs,ncovar", [(None, None), (200, 8), (60, 5)])
def test_full_sweep_matches_oracle(nsubs, ncovar, xtz_mode):
    """One bsg_sweep: the colour-1 kernel also forms the pair sums and draws Lambda."""
    print(f"case full sweep N={nsubs} P={ncovar} xtz_mode={xtz_mode}", flush=True)
    d, lay, L, S, O, sc0 = problem("c1", nsubs, ncovar, xtz_mode=xtz_mode)
    P = S.P
    sweep = 5
    S.set_spat_coef(sc0)
    S.set_sweep_index(sweep)
    prec0 = (np.eye(P) * 1.5 + 0.2).astype(np.float32)
    S.set_prec(prec0)
    S.sweep(1)
    S.synchronize()
    z_gpu, m_gpu, b_gpu = S.get_z(), S.get_xtz(), S.get_spat_coef()
    assert rel_err(z_gpu, O.update_z(sc0, sweep)) < 1e-5
    scale = np.abs(O.covar).astype(np.float64) @ np.abs(z_gpu.astype(np.float64))
    assert np.max(np.abs(m_gpu - O.xtz(z_gpu)) / scale) < (2e-5 if xtz_mode == 1 else 1e-5)
    ref = O.update_coef(0, m_gpu.astype(np.float64), prec0, sc0.astype(np.float64), sweep)
    ref = O.update_coef(1, m_gpu.astype(np.float64), prec0, ref, sweep)
    assert rel_err(b_gpu, ref) < 1e-5
    t_gpu = S.get_pair_sums()
    t_orc = O.pair_sums(b_gpu)
    assert np.max(np.abs(t_gpu - t_orc) / np.abs(t_orc).max()) < 1e-12
    lam_orc = O.wishart(t_gpu, lay["totvox"] + 19, sweep)
    lam_gpu = S.get_prec()
    assert rel_err(lam_gpu, lam_orc, floor=np.abs(lam_orc).max()) < 1e-6
    assert S.sweep_index() == sweep + 1
    # a second sweep: the side-stream Wishart of the first must be ordered before it
    S.sweep(1)
    S.synchronize()
    assert S.sweep_index() == sweep + 2
    t2 = S.get_pair_sums()
    assert np.max(np.abs(t2 - O.pair_sums(S.get_spat_coef())) / np.abs(t2).max()) < 1e-12


def test_z_store_streaming():
    """Streaming stores (z_store 1) give the same Z and X'Z as plain stores."""
    print("case c1 z_store=1", flush=True)
    d, lay, L, S, O, sc0 = problem("c1", xtz_mode=2)
    S2 = Sampler(L, d["covar"], d["data"], seed=SEED, xtz_mode=2, z_store=1)
    for T in (S, S2):
        T.set_spat_coef(sc0)
        T.set_sweep_index(3)
        T.update_latent()
    assert np.array_equal(S.get_z(), S2.get_z())
    assert np.array_equal(S.get_xtz(), S2.get_xtz())


# (N, P) cases span every lane split of the voxel-per-lane kernel:
# N=5 one lane per voxel, N=37/50 two, N=300 four, N=600 eight
@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("nsubs,ncovar,nchains", [(5, 2, 1), (37, 8, 1), (None, None, 2), (300, 4, 1),
                                                  (600, 3, 1), (140, 1, 1)])
def test_z_free_fused_sweep(nsubs, ncovar, nchains, method):
    """Without Z (z_store 2) X'Z is formed from draws held in registers: it
    must match X' times the oracle's own draws (same Philox counters), and the
    coefficient updates that consume it must match the oracle's."""
    check_z_free(nsubs, ncovar, nchains, method)


# Lane splits of the inversion path: G = 4 forms the means on the matrix
# cores with one MFMA for P = 3, 4, two for P = 5..8 and none for P <= 2; one
# chain at G = 4 also splits the leftover tiles (and the latent launch carries
# the previous sweep's Wishart draw).
@pytest.mark.parametrize("nsubs,ncovar,g,nchains", [(300, 6, 1, 1), (300, 6, 2, 1), (300, 6, 4, 2), (300, 6, 8, 1),
                                                    (200, 8, 4, 1), (200, 8, 4, 2), (90, 2, 4, 2), (57, 5, 4, 1),
                                                    (13, 3, 4, 2), (150, 7, 4, 1)])
def test_z_free_every_lane_split(nsubs, ncovar, g, nchains):
    check_z_free(nsubs, ncovar, nchains, 0, g)


# X'Z as BF16 hi + lo products on the matrix cores inside the latent draws
# (xtz_mode 3, G = 4): P = 1..8, leftover tiles in pieces (N >= 57) or none
# (N = 13), several chains (whole tiles), one lane column per covariate
@pytest.mark.parametrize("nsubs,ncovar,nchains", [(300, 4, 1), (200, 8, 1), (200, 8, 2), (57, 5, 1), (13, 3, 2),
                                                  (90, 2, 1), (150, 1, 1), (150, 7, 1), (37, 6, 1), (2, 2, 1),
                                                  (4, 8, 2)])
def test_bf16_mfma_xtz_fused_into_the_draws(nsubs, ncovar, nchains):
    check_z_free(nsubs, ncovar, nchains, 0, 0, xtz_mode=3)


def test_bf16_mfma_fused_form_limits():
    """xtz_mode 3 needs the Z-free form at four lanes per voxel with
    inversion draws and no WM image: other settings are refused."""
    d = synth.generate("c1", nsubs=60, ncovar=4)
    L = Layout(d["mask"], *d["grid"])
    for kw in (dict(z_store=0), dict(z_store=1), dict(z_store=2, latent_method=1)):
        with pytest.raises(Exception):
            Sampler(L, d["covar"], d["data"], seed=1, xtz_mode=3, **kw)
    with pytest.raises(Exception):
        Sampler(L, d["covar"], d["data"], wm=np.ones(L.totvox, np.float32), seed=1, xtz_mode=3, z_store=2)
    S = Sampler(L, d["covar"], d["data"], seed=1, xtz_mode=3, z_store=2)
    assert S.lane_split == 4
    with pytest.raises(Exception):
        S.lane_split = 8
    S.lane_split = 4
    g, ms = S.autotune(2)
    assert g == 4 and [m for m in ms if m >= 0] == [ms[2]]


def test_split_tail_single_steps_and_full_size():
    """Four lanes per voxel split a one-chain launch's leftover 64-voxel tiles
    into pieces whose partial X'Z sums the next coefficient kernel (or
    get_xtz) adds. Small mask: fewer tiles than compute units, so every tile
    is split; the X'Z read after the latent step alone against the oracle,
    then X'Z set by the host (which rewrites the pieces) and both colours'
    updates against the oracle on it. Full C2 size: whole tiles plus split
    leftovers, against the unsplit eight-lanes-per-voxel form."""
    print("case split tail: c1 N=150 P=4 steps; c2 full size G=4 vs G=8", flush=True)
    d = synth.generate("c1", nsubs=150, ncovar=4)
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    L = Layout(d["mask"], nx, ny, nz)
    S = Sampler(L, d["covar"], d["data"], seed=SEED, xtz_mode=2, z_store=2)
    assert S.lane_split == 4
    O = oracle.Oracle(lay, d["covar"], d["data"], seed=SEED)
    rng = np.random.default_rng(12)
    prec0 = (np.eye(4) * 1.5 + 0.2).astype(np.float32)
    sc0 = spat_coef_from_ranks(d["beta_true"] + 0.1 * rng.standard_normal(d["beta_true"].shape)
                               .astype(np.float32), lay)
    sweep = 6
    S.set_sweep_index(sweep)
    S.set_spat_coef(sc0)
    S.set_prec(prec0)
    S.update_latent()
    X = O.covar.astype(np.float64)
    z_orc = O.update_z(sc0, sweep)
    scale = np.abs(X) @ np.abs(z_orc)
    assert np.max(np.abs(S.get_xtz() - O.xtz(z_orc.astype(np.float32))) / scale) < 2e-5
    m_set = (O.xtz(z_orc) * (1.0 + 1e-3)).astype(np.float32)  # host values differ from the pieces
    S.set_xtz(m_set)
    assert np.array_equal(S.get_xtz(), m_set)
    S.update_coef(0)
    S.update_coef(1)
    assert np.array_equal(S.get_xtz(), m_set)
    ref = O.update_coef(0, m_set.astype(np.float64), prec0, sc0.astype(np.float64), sweep)
    ref = O.update_coef(1, m_set.astype(np.float64), prec0, ref, sweep)
    assert rel_err(S.get_spat_coef(), ref) < 1e-5

    d = synth.generate("c2")
    L = Layout(d["mask"], *d["grid"])
    A = Sampler(L, d["covar"], d["data"], seed=SEED, xtz_mode=2, z_store=2)
    B = Sampler(L, d["covar"], d["data"], seed=SEED, xtz_mode=2, z_store=2)
    B.lane_split = 8
    assert A.lane_split == 4
    sc = spat_coef_from_ranks(d["beta_true"], oracle.layout(d["mask"], *d["grid"]))
    for S2 in (A, B):
        S2.set_spat_coef(sc)
        S2.set_sweep_index(3)
        S2.update_latent()
    ma, mb = A.get_xtz(), B.get_xtz()
    # same Philox draws; the means (MFMA vs FMA) and the X'Z order differ in FP32 rounding
    assert np.max(np.abs(ma - mb) / (np.abs(mb) + 10.0)) < 1e-5
    A.sweep(2)
    B.sweep(2)
    A.synchronize()
    B.synchronize()
    assert np.all(np.isfinite(A.get_spat_coef())) and np.all(np.isfinite(A.get_prec()))


def check_z_free(nsubs, ncovar, nchains, method, g=0, xtz_mode=2):
    print(f"case c1 z_store=2 xtz_mode={xtz_mode} method={method} N={nsubs} P={ncovar} chains={nchains} G={g}",
          flush=True)
    d = synth.generate("c1", nsubs=nsubs, ncovar=ncovar)
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    L = Layout(d["mask"], nx, ny, nz)
    S = Sampler(L, d["covar"], d["data"], seed=SEED, nchains=nchains, xtz_mode=xtz_mode, z_store=2,
                latent_method=method)
    if g:
        S.lane_split = g
        assert S.lane_split == g
    O = oracle.Oracle(lay, d["covar"], d["data"], seed=SEED)
    P = S.P
    rng = np.random.default_rng(11)
    sweep = 3
    prec0 = (np.eye(P) * 1.5 + 0.2).astype(np.float32)
    S.set_sweep_index(sweep)
    scs = []
    for c in range(nchains):
        beta = d["beta_true"] + 0.1 * rng.standard_normal(d["beta_true"].shape).astype(np.float32)
        scs.append(spat_coef_from_ranks(beta, lay))
        S.set_spat_coef(scs[c], chain=c)
        S.set_prec(prec0, chain=c)
    S.sweep(1)
    S.synchronize()
    T = None
    if xtz_mode == 3:
        # the same sweep with FP32 X'Z: the draws are the same, the X'Z sums are not
        T = Sampler(L, d["covar"], d["data"], seed=SEED, nchains=nchains, xtz_mode=2, z_store=2)
        T.set_sweep_index(sweep)
        for c in range(nchains):
            T.set_spat_coef(scs[c], chain=c)
            T.set_prec(prec0, chain=c)
        T.sweep(1)
        T.synchronize()
    for c in range(nchains):
        O.set_chain(c)
        z_orc = O.update_z(scs[c], sweep, method=method)
        m_gpu = S.get_xtz(c)
        if T is not None:
            assert not np.array_equal(m_gpu, T.get_xtz(c)), c  # the BF16 MFMA path ran
        # floor 1, as the draws' own bar: with a handful of subjects a covariate may reach one
        # subject only, whose draw can lie near 0, where the FP32 draw error is absolute
        scale = np.maximum(np.abs(O.covar).astype(np.float64) @ np.abs(z_orc), 1.0)
        # the GPU draws are FP32 (1e-5 of the oracle's) and summed in FP32 (xtz_mode 3: as BF16 hi + lo
        # products, each operand within 2^-18 of its value)
        assert np.max(np.abs(m_gpu - O.xtz(z_orc.astype(np.float32))) / scale) < 2e-5
        ref = O.update_coef(0, m_gpu.astype(np.float64), prec0, scs[c].astype(np.float64), sweep)
        ref = O.update_coef(1, m_gpu.astype(np.float64), prec0, ref, sweep)
        assert rel_err(S.get_spat_coef(c), ref) < 1e-5
    assert S.fallback_count() == 0
    with pytest.raises(Exception):
        S.get_z()


def test_full_size_invariants_and_determinism():
    print("case c2 full size, 3 sweeps x 2 contexts", flush=True)
    d = synth.generate("c2")
    nx, ny, nz = d["grid"]
    L = Layout(d["mask"], nx, ny, nz)
    runs = []
    for _ in range(2):
        S = Sampler(L, d["covar"], d["data"], seed=SEED)
        S.sweep(3)
        S.synchronize()
        runs.append((S.get_spat_coef(), S.get_prec(), S.get_z()))
        S.close()
    (b0, l0, z0), (b1, l1, z1) = runs
    assert np.array_equal(b0, b1) and np.array_equal(l0, l1) and np.array_equal(z0, z1)
    y = d["data"] == 1
    assert np.all(z0[y] > 0) and np.all(z0[~y] < 0)
    assert np.all(np.isfinite(b0)) and np.all(np.linalg.eigvalsh(l0.astype(np.float64)) > 0)


def test_chains_are_independent_streams():
    print("case c1, 3 chains", flush=True)
    d, lay, L, S, O, sc0 = problem("c1", nchains=3)
    S.sweep(2)
    S.synchronize()
    b = [S.get_spat_coef(c) for c in range(3)]
    assert not np.array_equal(b[0], b[1]) and not np.array_equal(b[1], b[2])
    # chain c of a 3-chain context equals chain c of the oracle stream: check chain 2's latent
    S2 = Sampler(L, d["covar"], d["data"], seed=SEED, nchains=3)
    S2.set_spat_coef(sc0, chain=2)
    S2.update_latent()
    O.set_chain(2)
    assert rel_err(S2.get_z(chain=2), O.update_z(sc0, 0)) < 1e-5


@pytest.mark.parametrize("z_store,xtz_mode,method,nsubs", [(0, 2, 0, None), (0, 0, 0, None), (2, 2, 0, None),
                                                           (2, 2, 1, None), (2, 2, 0, 150)])
def test_wm_coefficient_sweep(z_store, xtz_mode, method, nsubs):
    """With a WM image the sweep draws b between the latent draw and the
    coefficients (mcmc.cpp:344-356): Z uses the old b, the coefficients use
    X'(Z - b_new WM). Each step is checked against the oracle. N = 150 puts
    the Z-free form at four lanes per voxel (means on the matrix cores)."""
    print(f"case c1 WM z_store={z_store} xtz_mode={xtz_mode} method={method} N={nsubs}", flush=True)
    d = synth.generate("c1", nsubs=nsubs)
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    L = Layout(d["mask"], nx, ny, nz)
    rng = np.random.default_rng(5)
    wm = rng.uniform(0.0, 1.0, lay["totvox"]).astype(np.float32)
    S = Sampler(L, d["covar"], d["data"], wm=wm, seed=SEED, xtz_mode=xtz_mode, z_store=z_store,
                latent_method=method)
    O = oracle.Oracle(lay, d["covar"], d["data"], wm=wm, seed=SEED)
    P = S.P
    sweep, b0 = 7, 0.3
    sc0 = spat_coef_from_ranks(d["beta_true"], lay)
    prec0 = (np.eye(P) * 1.5 + 0.2).astype(np.float32)
    S.set_spat_coef(sc0)
    S.set_prec(prec0)
    S.set_wm_coef(b0)
    S.set_sweep_index(sweep)
    S.sweep(1)
    S.synchronize()
    z_orc = O.update_z(sc0, sweep, b=b0, method=method)
    if z_store == 0:
        z_gpu = S.get_z()
        assert rel_err(z_gpu, z_orc) < 1e-5
        z_used = z_gpu.astype(np.float64)
    else:
        z_used = z_orc  # Z is not kept; the oracle's draws carry the same counters
    s_orc, w_orc = O.wm_stat(sc0, z_used)
    b_orc = O.wm_draw(s_orc, w_orc, sweep)
    b_gpu = S.get_wm_coef()
    # FP32 draws and per-voxel sums against FP64: the mean S/W moves by < 1e-6
    assert abs(b_gpu - b_orc) < 1e-6 + 1e-5 * abs(b_orc), (b_gpu, b_orc)
    m_ref = O.xtz(z_used.astype(np.float32), b_gpu)
    ref = O.update_coef(0, m_ref, prec0, sc0.astype(np.float64), sweep)
    ref = O.update_coef(1, m_ref, prec0, ref, sweep)
    assert rel_err(S.get_spat_coef(), ref) < 2e-5
    # the statistic itself: b ~ N(S/W, 1/W) about the oracle's mean
    assert abs(b_gpu - s_orc / w_orc) < 6.0 / np.sqrt(w_orc)


@pytest.mark.parametrize("nsubs,ncovar,nchains", [(150, 4, 2), (90, 6, 1), (300, 3, 1), (200, 8, 1), (100, 7, 2)])
def test_wishart_carried_by_next_latent_launch(nsubs, ncovar, nchains):
    """At four lanes per voxel block 0 of the Z-free latent launch of sweep t+1
    draws the CAR precision of sweep t, and the last sweep of a call draws it in
    a launch of its own. After every call Lambda must be the oracle's Wishart of
    the GPU's own pair sums, and sweeps split over several calls must equal the
    same sweeps in one call bit for bit."""
    print(f"case c1 carried Wishart N={nsubs} P={ncovar} chains={nchains}", flush=True)
    d = synth.generate("c1", nsubs=nsubs, ncovar=ncovar)
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    L = Layout(d["mask"], nx, ny, nz)
    mk = lambda: Sampler(L, d["covar"], d["data"], seed=SEED, nchains=nchains, xtz_mode=2, z_store=2)
    A, B = mk(), mk()
    assert A.lane_split == 4
    O = oracle.Oracle(lay, d["covar"], d["data"], seed=SEED)
    for n in (1, 3, 4):
        A.sweep(n)
        A.synchronize()
        for c in range(nchains):
            O.set_chain(c)
            t = A.get_pair_sums(c)
            assert np.max(np.abs(t - O.pair_sums(A.get_spat_coef(c))) / np.abs(t).max()) < 1e-12
            lam_orc = O.wishart(t, lay["totvox"] + 19, A.sweep_index() - 1)
            assert rel_err(A.get_prec(c), lam_orc, floor=np.abs(lam_orc).max()) < 1e-6
    B.sweep(8)
    B.synchronize()
    assert A.sweep_index() == B.sweep_index()
    for c in range(nchains):
        assert np.array_equal(A.get_spat_coef(c), B.get_spat_coef(c))
        assert np.array_equal(A.get_prec(c), B.get_prec(c))


@pytest.mark.parametrize("z_store,xtz_mode,wm,nsubs", [(2, 2, False, None), (0, 1, True, None), (2, 2, False, 150),
                                                       (2, 2, True, 150), (2, 3, False, 150)])
def test_graph_replays_equal_direct_sweeps(z_store, xtz_mode, wm, nsubs):
    """Sweeps replayed from a captured HIP graph (device sweep base advanced by
    the graph's last node) give bit-identical state to direct launches, over
    several replays with direct sweeps and host reads between them. N = 150
    runs at four lanes per voxel, where each latent launch carries the previous
    sweep's Wishart draw."""
    print(f"case c1 graph z_store={z_store} xtz_mode={xtz_mode} wm={wm} N={nsubs}", flush=True)
    d = synth.generate("c1", nsubs=nsubs)
    nx, ny, nz = d["grid"]
    L = Layout(d["mask"], nx, ny, nz)
    w = np.random.default_rng(1).uniform(0, 1, L.totvox).astype(np.float32) if wm else None
    A = Sampler(L, d["covar"], d["data"], wm=w, seed=SEED, nchains=2, xtz_mode=xtz_mode, z_store=z_store)
    G = Sampler(L, d["covar"], d["data"], wm=w, seed=SEED, nchains=2, xtz_mode=xtz_mode, z_store=z_store)
    for step in (11, 3, 8, 16):  # 11 = one group of 8 + 3 direct; then a direct batch, then replays again
        A.sweep(step)
        if step == 3:
            G.sweep(step)
        else:
            G.sweep_graph(step, 8)
        A.synchronize()
        G.synchronize()
        assert G.sweep_index() == A.sweep_index()
        for c in range(2):
            assert np.array_equal(G.get_spat_coef(c), A.get_spat_coef(c))
            assert np.array_equal(G.get_prec(c), A.get_prec(c))
            assert G.get_wm_coef(c) == A.get_wm_coef(c)


def test_graph_recaptured_after_lane_split_change():
    """A captured sweep graph holds the lane split by value: after the split
    changes (here 2 -> 8 -> 1, none of them overlapped, so the graph's form key
    stays the same), graph sweeps must still equal direct sweeps."""
    print("case c1 N=150 graph after lane-split changes", flush=True)
    d = synth.generate("c1", nsubs=150)
    L = Layout(d["mask"], *d["grid"])
    A = Sampler(L, d["covar"], d["data"], seed=SEED, xtz_mode=2, z_store=2)
    G = Sampler(L, d["covar"], d["data"], seed=SEED, xtz_mode=2, z_store=2)
    for g in (2, 8, 1):
        A.lane_split = g
        G.lane_split = g
        assert not G.overlap
        A.sweep(8)
        G.sweep_graph(8, 8)
        A.synchronize()
        G.synchronize()
        assert np.array_equal(A.get_spat_coef(), G.get_spat_coef()), g
        assert np.array_equal(A.get_prec(), G.get_prec()), g
