"""The batched configurations at their own shapes.

C4: the full-brain C2 data (V = 230,906, N = 250, P = 4) as 32 chains in one
context. The overlapped sweep (two launches per sweep, every chain in each)
must equal the serial order bit for bit, every chain must stay finite with a
positive definite CAR precision, and chain 31's X'Z must be X' times the
oracle's own chain-31 draws (its own Philox stream).

C5: the twelve uneven analyses c5_0 ... c5_11 (ellipsoid masks of 20,031 to
398,905 voxels, N from 40 to 1,000, P from 2 to 8), 4 chains each, in one
batch, with X'Z as BF16 hi + lo products on the matrix cores (the
configuration's form) and in FP32: each member's sweeps must equal the member
swept alone, bit for bit.

Batched maps: 32 chains of the C1 data in one context, run with the reference
CPU path's semantics, with X'Z in FP32 and in the mixed-precision form (BF16
hi + lo products on the matrix cores); their pooled posterior means must agree with a long
chain of the reference's own CPU path (tests/golden/ref_chain_c1.npz) within
4 combined Monte-Carlo standard errors, up to the exceedances a Student-t
null allows (the MCSEs are batch-means estimates, so |z| > 4 is more likely
than under a normal: the bound is the binomial 1e-4 upper quantile of the
count at P(|t_nu| > 4), nu = reference batches - 1).
"""
import os

import numpy as np
import pytest

import oracle
from bsglmm_amd import Batch, Layout, Sampler, synth

pytestmark = pytest.mark.gpu
SEED = 0xC4C5C4C5
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def spat_coef_from_ranks(beta, lay):
    hi, hs = lay["host_idx"], lay["host_idx_sc"]
    ids = np.flatnonzero(hi >= 0)
    sc = np.zeros((beta.shape[0], lay["totvoxp"]), np.float32)
    sc[:, hs[ids]] = beta[:, hi[ids]]
    return sc


def exceedance_bound(n, nu, z=4.0, q=1e-4):
    """Smallest k with P(count > k) < q when each of n comparisons exceeds |z|
    with probability P(|t_nu| > z)."""
    from scipy import stats
    p0 = 2.0 * stats.t.sf(z, nu)
    return int(stats.binom.ppf(1.0 - q, n, p0)), n * p0


def test_c4_32_chains_overlapped_equals_serial():
    print("case C4: c2 data, 32 chains, 3 sweeps overlapped vs serial", flush=True)
    d = synth.generate("c2")
    nx, ny, nz = d["grid"]
    L = Layout(d["mask"], nx, ny, nz)
    C = 32
    A = Sampler(L, d["covar"], d["data"], seed=SEED, nchains=C, xtz_mode=2, z_store=2)
    B = Sampler(L, d["covar"], d["data"], seed=SEED, nchains=C, xtz_mode=2, z_store=2)
    B.overlap = False
    assert A.overlap and not B.overlap and A.lane_split == 4
    for X in (A, B):
        X.sweep(2)
        X.synchronize()
    sc_before = A.get_spat_coef(31)
    for X in (A, B):
        X.sweep(1)
        X.synchronize()
    assert A.sweep_index() == B.sweep_index() == 3
    for c in range(C):
        ba, la = A.get_spat_coef(c), A.get_prec(c)
        assert np.array_equal(ba, B.get_spat_coef(c)), c
        assert np.array_equal(la, B.get_prec(c)), c
        assert np.all(np.isfinite(ba))
        assert np.all(np.linalg.eigvalsh(la.astype(np.float64)) > 0)
    # chain 31's X'Z of sweep 2, from its own coefficients after sweep 1
    lay = oracle.layout(d["mask"], nx, ny, nz)
    O = oracle.Oracle(lay, d["covar"], d["data"], seed=SEED)
    O.set_chain(31)
    z = O.update_z(sc_before, 2)
    m = A.get_xtz(31)
    scale = np.abs(O.covar).astype(np.float64) @ np.abs(z)
    err = float(np.max(np.abs(m - O.xtz(z.astype(np.float32))) / scale))
    print(f"chain 31 X'Z error / sum|x||z|: {err:.2e}", flush=True)
    assert err < 2e-5
    # distinct Philox streams per chain
    assert not np.array_equal(A.get_spat_coef(0), A.get_spat_coef(31))


# xtz_mode 3 is the configuration's form (X'Z as BF16 hi + lo products on the
# matrix cores, FP32 accumulation, here inside the draws); 2 the FP32 one
@pytest.mark.parametrize("xtz_mode", [3, 2])
def test_c5_twelve_analysis_batch_equals_members_alone(xtz_mode):
    print(f"case C5: c5_0..c5_11, 4 chains each, xtz_mode {xtz_mode}, 2 batch sweeps vs each member alone",
          flush=True)
    solo, grouped = [], []
    for k in range(12):
        d = synth.generate(f"c5_{k}")
        L = Layout(d["mask"], *d["grid"])
        for lst in (solo, grouped):
            lst.append(Sampler(L, d["covar"], d["data"], seed=SEED + k, nchains=4, xtz_mode=xtz_mode, z_store=2))
        del d
    Ps = sorted({S.P for S in grouped})
    assert Ps == [2, 3, 4, 5, 6, 7, 8]
    B = Batch(grouped)
    assert not B.one_launch_per_phase  # 3.1e9 draws per sweep: one launch per phase and P
    # two sweeps one launch per phase and P, then two in one launch per phase
    for mode, total in ((2, 2), (1, 4)):
        B.set_launch_mode(mode)
        B.sweep(2)
        for S in solo:
            S.sweep(2)
        for k, (S, T) in enumerate(zip(solo, grouped)):
            S.synchronize()
            T.synchronize()
            assert S.sweep_index() == T.sweep_index() == total
            for c in range(4):
                assert np.array_equal(S.get_spat_coef(c), T.get_spat_coef(c)), (mode, k, c)
                assert np.array_equal(S.get_prec(c), T.get_prec(c)), (mode, k, c)
                assert np.all(np.isfinite(T.get_spat_coef(c)))
    B.close()
    for S in solo + grouped:
        S.close()


# the batched chains in FP32 (xtz_mode 2) and in the mixed-precision form
# (xtz_mode 3: X'Z as BF16 hi + lo products on the matrix cores)
@pytest.mark.parametrize("xtz_mode", [2, 3])
def test_batched_chains_pooled_maps_match_reference_cpu_chain(xtz_mode):
    f = np.load(os.path.join(GOLD, "ref_chain_c1.npz"))
    C, burnin, draws, nb = 32, 1000, 2000, 20
    print(f"case batched maps: c1, {C} chains x ({burnin}+{draws}) sweeps, xtz_mode {xtz_mode}, vs the reference "
          f"CPU chain", flush=True)
    d = synth.generate("c1")
    L = Layout(d["mask"], *d["grid"])
    S = Sampler(L, d["covar"], d["data"], seed=0xBA7C4, nchains=C, xtz_mode=xtz_mode, z_store=2,
                df_offset=int(f["df_offset"]), trunc_eps=float(f["trunc_eps"]))
    assert S.overlap
    S.sweep(burnin)
    S.accum_reset(ntypes=1, nbatches=nb, batch_len=draws // nb)
    S.sweep_accumulate(draws, 1)
    S.synchronize()
    means, mcses = [], []
    for c in range(C):
        o = S.accum_get(c)
        assert o["ndraws"] == draws
        means.append(np.vstack([o["mean"], o["prb"]]))
        mcses.append(o["mcse"])
    pooled = np.mean(means, axis=0)
    pooled_mcse = np.sqrt(np.sum(np.square(mcses), axis=0)) / C
    z = (pooled - f["mean"]) / np.sqrt(pooled_mcse ** 2 + f["mcse"].astype(np.float64) ** 2)
    n = z.size
    count = int(np.sum(np.abs(z) > 4.0))
    bound, expect = exceedance_bound(n, int(f["batches"]) - 1)
    print(f"pooled {C}-chain maps: {count} of {n} beyond 4 MCSE (null expectation {expect:.1f}, bound {bound}), "
          f"median |z| {np.median(np.abs(z)):.3f}", flush=True)
    assert count <= bound
    assert np.median(np.abs(z)) < 1.2
    # the chains mixed: between-chain spread of the means agrees with the within-chain MCSE
    spread = np.std(means, axis=0, ddof=1)
    ratio = np.median(spread / np.maximum(np.sqrt(np.mean(np.square(mcses), axis=0)), 1e-12))
    print(f"median between-chain sd / within-chain MCSE: {ratio:.3f}", flush=True)
    assert 0.7 < ratio < 1.4
