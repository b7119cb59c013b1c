"""Overlapped sweep (k_overlap): each colour's coefficient update runs in the
launch that draws the other colour's latent variates, and the CAR precision
is drawn inside the launch whose colour-0 blocks read it. It must reproduce
the serial sweep bit for bit: the same Philox counters, the same colour tiles
and pieces, the same sums in the same order (mcmc.cpp:311-394 order of
updates). Checked on the coefficients, X'Z, Lambda, the pair sums, the
posterior accumulators and across call boundaries and HIP-graph replays.
"""
import numpy as np
import pytest

import oracle
from bsglmm_amd import Layout, Sampler, lib, synth

pytestmark = pytest.mark.gpu

SEED = 0x1234ABCD5678EF01


def spat_coef_from_ranks(beta, lay):
    hi, hs = lay["host_idx"], lay["host_idx_sc"]
    ids = np.flatnonzero(hi >= 0)
    sc = np.zeros((beta.shape[0], lay["totvoxp"]), np.float32)
    sc[:, hs[ids]] = beta[:, hi[ids]]
    return sc


def pair(name, nsubs, ncovar, g=0, nchains=1, wm=False, xtz_mode=2):
    kw = {} if name in ("c2", "c3") else dict(nsubs=nsubs, ncovar=ncovar)
    d = synth.generate(name, **kw)
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    L = Layout(d["mask"], nx, ny, nz)
    w = None
    if wm:
        w = (0.5 + np.random.default_rng(3).random(d["data"].shape[1])).astype(np.float32)
    out = []
    for on in (True, False):
        S = Sampler(L, d["covar"], d["data"], wm=w, seed=SEED, nchains=nchains, xtz_mode=xtz_mode, z_store=2)
        if g:
            S.lane_split = g
        S.overlap = on
        out.append(S)
    rng = np.random.default_rng(21)
    sc = spat_coef_from_ranks(d["beta_true"] + 0.05 * rng.standard_normal(d["beta_true"].shape)
                              .astype(np.float32), lay)
    for S in out:
        for c in range(nchains):
            S.set_spat_coef(sc, chain=c)
        S.set_sweep_index(7)
    return L, out[0], out[1]


def assert_same_state(A, B, nchains=1):
    A.synchronize()
    B.synchronize()
    for c in range(nchains):
        assert A.sweep_index(c) == B.sweep_index(c)
        assert np.array_equal(A.get_spat_coef(c), B.get_spat_coef(c))
        assert np.array_equal(A.get_xtz(c), B.get_xtz(c))
        assert np.array_equal(A.get_prec(c), B.get_prec(c))
        assert np.array_equal(A.get_pair_sums(c), B.get_pair_sums(c))


# (N, P, lane split): P = 1..8 (means on FMA for P <= 2, one MFMA for P = 3..4,
# two for P = 5..8); leftover tiles in 4 pieces at N >= 57, none at N = 13
# (one subject round); C2 at full size (13 and 12 leftover tiles in 16 pieces);
# the C3 shape (N = 1,000, P = 8) at full size
@pytest.mark.parametrize("name,nsubs,ncovar,g", [("c1", 300, 4, 0), ("c1", 200, 6, 0), ("c1", 57, 5, 4),
                                                 ("c1", 13, 3, 4), ("c1", 90, 2, 0), ("c1", 150, 1, 0),
                                                 ("c1", 120, 7, 0), ("c1", 200, 8, 0), ("c2", None, None, 0),
                                                 ("c3", None, None, 0)])
def test_overlapped_sweep_equals_serial(name, nsubs, ncovar, g):
    print(f"case overlap vs serial {name} N={nsubs} P={ncovar} G={g}", flush=True)
    L, A, B = pair(name, nsubs, ncovar, g)
    assert A.overlap and not B.overlap
    A.sweep(2)  # two calls: the first sweep of a call reads the Lambda the call before drew
    A.sweep(3)
    B.sweep(5)
    assert_same_state(A, B)
    assert A.fallback_count() == B.fallback_count()


# the X'Z of the draws as BF16 hi + lo products on the matrix cores (xtz_mode 3):
# one chain (colour tiles, pieces or none) and several (whole tiles), C3 at
# full size
@pytest.mark.parametrize("name,nsubs,ncovar,nchains", [("c1", 300, 4, 1), ("c1", 57, 5, 1), ("c1", 13, 3, 1),
                                                       ("c1", 90, 2, 1), ("c1", 200, 8, 1), ("c1", 300, 4, 2),
                                                       ("c1", 150, 7, 3), ("c3", None, None, 1)])
def test_bf16_fused_overlapped_sweep_equals_serial(name, nsubs, ncovar, nchains):
    print(f"case xtz_mode 3 overlap vs serial {name} N={nsubs} P={ncovar} chains={nchains}", flush=True)
    L, A, B = pair(name, nsubs, ncovar, nchains=nchains, xtz_mode=3)
    A.overlap = True
    assert A.overlap and not B.overlap
    A.sweep(2)
    A.sweep(3)
    B.sweep(5)
    assert_same_state(A, B, nchains)


def test_overlapped_accumulation_and_graph_equal_serial():
    """Save sweeps accumulate inside the overlapped colour updates; a graph of
    overlapped sweeps replays the same chain as direct serial sweeps."""
    print("case overlap accumulation + graph c1 N=300 P=4", flush=True)
    L, A, B = pair("c1", 300, 4)
    for S in (A, B):
        S.accum_reset(ntypes=2, nbatches=4, batch_len=2)
        S.sweep_accumulate(8, save_every=2)
    assert_same_state(A, B)
    ra, rb = A.accum_get(), B.accum_get()
    for k in ra:
        assert np.array_equal(np.asarray(ra[k]), np.asarray(rb[k]), equal_nan=True), k
    A.sweep_graph(9, per_graph=4)  # two replays of 4 and one direct sweep
    B.sweep(9)
    assert_same_state(A, B)
    A.sweep_graph(8, per_graph=4)
    B.sweep(8)
    assert_same_state(A, B)


# several chains: the grid's second dimension, whole tiles (no pieces), one
# Lambda draw and wait per chain; C2 with 3 chains at full size
@pytest.mark.parametrize("name,nsubs,ncovar,nchains", [("c1", 300, 4, 2), ("c1", 200, 6, 3), ("c1", 90, 2, 2),
                                                       ("c1", 100, 8, 2), ("c1", 60, 7, 3), ("c2", None, None, 3)])
def test_overlapped_sweep_equals_serial_several_chains(name, nsubs, ncovar, nchains):
    print(f"case overlap vs serial {name} N={nsubs} P={ncovar} chains={nchains}", flush=True)
    L, A, B = pair(name, nsubs, ncovar, nchains=nchains)
    assert A.overlap and not B.overlap
    for S in (A, B):
        S.accum_reset(ntypes=1, nbatches=2, batch_len=2)
    A.sweep_accumulate(1, save_every=1)
    A.sweep_accumulate(3, save_every=1)
    B.sweep_accumulate(4, save_every=1)
    assert_same_state(A, B, nchains)
    for c in range(nchains):
        ra, rb = A.accum_get(c), B.accum_get(c)
        for k in ra:
            assert np.array_equal(np.asarray(ra[k]), np.asarray(rb[k]), equal_nan=True), (c, k)
    # the chains are distinct streams
    assert not np.array_equal(A.get_spat_coef(0), A.get_spat_coef(1))


def test_bounded_lambda_wait_reports_and_recovers():
    """The colour-0 blocks' wait for the Lambda drawn beside them is bounded.
    Forced from the host (bsg_test_overlap_wait): the draw writes a tag the
    blocks never accept, every wait gives up after 8 polls, the kernels
    finish, and bsg_synchronize reports it once; with the hook cleared the
    context sweeps on."""
    print("case overlap bounded wait forced from the host", flush=True)
    L, A, B = pair("c1", 300, 4)
    assert lib().bsg_test_overlap_wait(A._h, 8, 2) == 0
    A.sweep(3)  # phases 0 of the second and third sweeps wait
    with pytest.raises(Exception, match="past its bound"):
        A.synchronize()
    assert lib().bsg_test_overlap_wait(A._h, 0, 0) == 0
    A.sweep(2)
    A.synchronize()
    assert np.all(np.isfinite(A.get_spat_coef()))


def test_overlap_off_where_unsupported():
    """A WM image or a lane split other than four run the serial order."""
    print("case overlap unsupported forms", flush=True)
    for kw in (dict(wm=True),):
        L, A, B = pair("c1", 300, 4, **kw)
        assert not A.overlap
    L, A, B = pair("c1", 300, 4, g=8)
    assert not A.overlap
    A.lane_split = 4
    assert A.overlap


def test_prepared_graph_runs_no_sweep_and_replays_the_chain():
    """bsg_sweep_graph_prepare captures and uploads the graph without running
    it: the sweep index and state are unchanged; the replays that follow give
    the chain of direct serial sweeps."""
    print("case prepared graph c1 N=300 P=4", flush=True)
    L, A, B = pair("c1", 300, 4)
    before = A.get_spat_coef()
    A.prepare_graph(4)
    A.synchronize()
    assert A.sweep_index() == B.sweep_index()
    assert np.array_equal(A.get_spat_coef(), before)
    A.sweep_graph(8, per_graph=4)
    B.sweep(8)
    assert_same_state(A, B)
