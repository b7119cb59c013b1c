"""GPU chain-level checks: the device posterior accumulator, and posterior
means of a long GPU chain against a long chain of the reference's own CPU
path (tests/golden/ref_chain_c1.npz, made by tests/golden/make_ref_chain.py).

The two chains use different random streams (the reference's host KISS
versus Philox) and different exact truncated-normal samplers (Robert 1995
versus inversion), so agreement is statistical: every posterior mean must
agree within 4 combined Monte-Carlo standard errors, up to the exceedances a
Student-t null allows among the 28,224 comparisons (the binomial 1e-4 upper
quantile of the count at P(|t_49| > 4): 17, against 6.0 expected).
"""
import os

import numpy as np
import pytest

from bsglmm_amd import Layout, Sampler, synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def phi_as(x):
    """Abramowitz & Stegun 26.2.17, as the reference's ProbSDNorm (covar.cpp:327-348)."""
    t = 1.0 / (1.0 + 0.2316419 * np.abs(x))
    poly = t * (0.319381530 + t * (-0.356563782 + t * (1.781477937 + t * (-1.821255978 + t * 1.330274429))))
    tail = 0.39894228 * np.exp(-0.5 * x * x) * poly
    return np.where(x >= 0, 1.0 - tail, tail)


def exceedance_bound(n, nu, z=4.0, q=1e-4):
    """Smallest k with P(count > k) < q when each of n comparisons exceeds |z|
    with probability P(|t_nu| > z): the MCSEs are batch-means estimates with
    nu degrees of freedom, so |z| > 4 is likelier than under a normal."""
    from scipy import stats
    p0 = 2.0 * stats.t.sf(z, nu)
    return int(stats.binom.ppf(1.0 - q, n, p0)), n * p0


def rank_view(sc, L):
    e = L.export()
    hi, hs = e["host_idx"], e["host_idx_sc"]
    ids = np.flatnonzero(hi >= 0)
    col = np.empty(L.totvox, np.int64)
    col[hi[ids]] = hs[ids]
    return sc[:, col].astype(np.float64)


def test_accumulator_matches_host_snapshots():
    print("case c1 accumulator, 2 chains x 24 draws", flush=True)
    d = synth.generate("c1")
    nx, ny, nz = d["grid"]
    L = Layout(d["mask"], nx, ny, nz)
    S = Sampler(L, d["covar"], d["data"], seed=77, nchains=2, xtz_mode=2)
    S.sweep(5)
    S.accum_reset(ntypes=2, nbatches=4, batch_len=6)
    snaps = [[], []]
    for _ in range(24):
        S.sweep(1)
        S.accum_add()
        for c in range(2):
            snaps[c].append(rank_view(S.get_spat_coef(c), L))
    for c in range(2):
        out = S.accum_get(c)
        b = np.array(snaps[c])  # draws x P x V
        assert out["ndraws"] == 24
        assert np.allclose(out["mean"], b.mean(0), rtol=1e-6, atol=1e-6)
        assert np.allclose(out["var"], b.var(0, ddof=1), rtol=1e-4, atol=1e-7)
        prb = phi_as(b[:, :2])
        assert np.allclose(out["prb"], prb.mean(0), rtol=1e-6, atol=1e-7)
        bm = np.concatenate([b, prb], axis=1).reshape(4, 6, b.shape[1] + 2, -1).mean(1)
        mcse = bm.std(0, ddof=1) / 2.0
        assert np.allclose(out["mcse"], mcse, rtol=1e-3, atol=1e-6)
    # sweep_accumulate adds every save_every-th draw
    S.accum_reset(ntypes=1, nbatches=2, batch_len=3)
    S.sweep_accumulate(12, save_every=2)
    assert S.accum_get(0)["ndraws"] == 6


@pytest.mark.parametrize("nbatches,batch_len,ndraws", [(4, 3, 17), (10, 4, 10), (3, 5, 15)])
def test_mcse_uses_completed_batches_only(nbatches, batch_len, ndraws):
    """Batch-means MCSE over the completed batches of batch_len draws: draws
    past nbatches * batch_len (and those of an unfinished batch) add to the
    means but to no batch. Both the separate accumulation pass (k_accum) and
    the save sweeps' fused accumulation are checked against host batch means
    of the GPU's own draws."""
    print(f"case c1 MCSE nbatches={nbatches} batch_len={batch_len} draws={ndraws}", flush=True)
    d = synth.generate("c1", nsubs=150)
    L = Layout(d["mask"], *d["grid"])
    A, F = (Sampler(L, d["covar"], d["data"], seed=91, xtz_mode=2, z_store=2) for _ in range(2))
    for X in (A, F):
        X.sweep(3)
        X.accum_reset(ntypes=1, nbatches=nbatches, batch_len=batch_len)
    snaps = []
    for _ in range(ndraws):
        A.sweep(1)
        A.accum_add()
        snaps.append(rank_view(A.get_spat_coef(), L))
    F.sweep_accumulate(ndraws, 1)
    b = np.array(snaps)
    x = np.concatenate([b, phi_as(b[:, :1])], axis=1)  # draws x (P + 1) x V
    nb = min(ndraws // batch_len, nbatches)
    for X in (A, F):
        out = X.accum_get(0)
        assert out["ndraws"] == ndraws
        assert np.allclose(out["mean"], b.mean(0), rtol=1e-6, atol=1e-6)
        if nb > 1:
            bm = x[:nb * batch_len].reshape(nb, batch_len, *x.shape[1:]).mean(1)
            mcse = bm.std(0, ddof=1) / np.sqrt(nb)
            assert np.allclose(out["mcse"], mcse, rtol=1e-3, atol=1e-6)
        else:
            assert np.all(out["mcse"] == 0.0)


@pytest.mark.parametrize("z_store,method,g", [(0, 0, 0), (2, 1, 0), (2, 0, 4)])
def test_posterior_means_match_reference_cpu_chain(z_store, method, g):
    """z_store 0: every draw by inversion (k_latent); z_store 2 with method 1:
    Box-Muller candidates with the exact fallback (k_latent_vl); z_store 2
    with method 0 at four lanes per voxel: the overlapped sweep (k_overlap,
    colour tiles and pieces, the Lambda draw inside its first launch)."""
    f = np.load(os.path.join(GOLD, "ref_chain_c1.npz"))
    burnin, draws, nb = int(f["burnin"]), int(f["draws"]), int(f["batches"])
    print(f"case c1 chain {burnin}+{draws} sweeps vs reference CPU chain, z_store={z_store} method={method} G={g}",
          flush=True)
    d = synth.generate("c1")
    nx, ny, nz = d["grid"]
    L = Layout(d["mask"], nx, ny, nz)
    P = d["covar"].shape[0]
    # the reference CPU path's semantics: df = TOTVOX + P, truncation at 0
    S = Sampler(L, d["covar"], d["data"], seed=0xC0FFEE, xtz_mode=2, z_store=z_store, latent_method=method,
                df_offset=int(f["df_offset"]), trunc_eps=float(f["trunc_eps"]))
    if g:
        S.lane_split = g
        assert S.overlap
    S.sweep(burnin)
    S.accum_reset(ntypes=1, nbatches=nb, batch_len=draws // nb)
    S.sweep_accumulate(draws, 1)
    S.synchronize()
    out = S.accum_get(0)
    ours = np.vstack([out["mean"], out["prb"]])
    z = (ours - f["mean"]) / np.sqrt(out["mcse"] ** 2 + f["mcse"].astype(np.float64) ** 2)
    count = int(np.sum(np.abs(z) > 4.0))
    bound, expect = exceedance_bound(z.size, nb - 1)
    print(f"{count} of {z.size} beyond 4 MCSE (null expectation {expect:.1f}, bound {bound}), "
          f"median |z| {np.median(np.abs(z)):.3f}", flush=True)
    assert count <= bound, count
    assert np.median(np.abs(z)) < 1.2  # about 0.67 for a standard normal z
    assert S.fallback_count() == 0
    assert out["mean"].shape == (P, L.totvox)
