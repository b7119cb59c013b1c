"""Batches of independent analyses in one launch per phase, and the lane-split
autotuner.

Every kernel body of a batched launch is the one a single analysis runs, with
the same Philox counters, so a batch must reproduce each member's own sweeps
bit for bit (coefficients, CAR precision, posterior accumulators).
"""
import numpy as np
import pytest

from bsglmm_amd import Batch, Layout, Sampler, synth

pytestmark = pytest.mark.gpu
SEED = 0xBA7C4ED


def members(seed_shift=0):
    """Four analyses with different masks, N and chain counts; three covariate
    counts, so three streams, one of them shared by two analyses (P = 3)."""
    out = []
    for (name, n, p, chains) in (("c1", 50, 3, 2), ("c1", 7, 2, 1), ("c5_0", 120, 5, 3), ("c5_1", 30, 3, 1)):
        d = synth.generate(name, nsubs=n, ncovar=p)
        L = Layout(d["mask"], *d["grid"])
        out.append((d, L, chains))
    return out


def make(ms, xtz=(2,)):
    """Z-free samplers; member k forms X'Z in mode xtz[k % len(xtz)] (2: FP32
    FMA, 3: BF16 hi + lo on the matrix cores)."""
    return [Sampler(L, d["covar"], d["data"], seed=SEED + k, nchains=ch, xtz_mode=xtz[k % len(xtz)], z_store=2)
            for k, (d, L, ch) in enumerate(ms)]


@pytest.mark.parametrize("xtz", [(2,), (3,), (3, 2)])
@pytest.mark.parametrize("mode", [1, 2])
def test_batch_equals_members_run_alone(mode, xtz):
    """mode 1: one launch per phase over every covariate count (each block
    switches on its member's P); mode 2: one launch per phase and P, each P on
    its own stream. Members form X'Z on FP32 FMAs, as BF16 hi + lo products on
    the matrix cores, or both kinds in one batch."""
    print(f"case batch of 4 analyses (c1 N=50 P=3 x2, c1 N=7 P=2, c5_0 N=120 P=5 x3, c5_1 N=30 P=3), mode {mode}, "
          f"xtz_mode {xtz}", flush=True)
    ms = members()
    solo, grouped = make(ms, xtz), make(ms, xtz)
    B = Batch(grouped)
    B.set_launch_mode(mode)
    assert B.one_launch_per_phase == (mode == 1)
    for S in solo:
        S.sweep(7)
    B.sweep(7)
    for S in solo:
        S.accum_reset(ntypes=1, nbatches=4, batch_len=3)
        S.sweep_accumulate(12, 1)
    for S in grouped:
        S.accum_reset(ntypes=1, nbatches=4, batch_len=3)
    B.sweep_accumulate(12, 1)
    for S, T in zip(solo, grouped):
        S.synchronize()
        T.synchronize()
        assert S.sweep_index() == T.sweep_index() == 19
        for c in range(S.nchains):
            assert np.array_equal(S.get_spat_coef(c), T.get_spat_coef(c))
            assert np.array_equal(S.get_prec(c), T.get_prec(c))
            a, b = S.accum_get(c), T.accum_get(c)
            assert a["ndraws"] == b["ndraws"] == 12
            for k in ("mean", "var", "prb", "mcse"):
                assert np.array_equal(a[k], b[k]), k
    B.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_precision_form_equals_members_run_alone(mixed):
    """Members in the precision form (Z stored by the k_latent body, X'Z on
    the matrix cores with BF16 hi+lo by the k_xtz_mfma body): the batch runs
    each member's own blocks, so every member's coefficients, X'Z, CAR
    precision and accumulators equal its sweeps alone bit for bit; mixed with
    Z-free members in one batch as well (one launch per phase and P)."""
    print(f"case precision-form batch (c1 N=50 P=3 x2, c1 N=7 P=2, c5_0 N=120 P=5 x3, c5_1 N=30 P=8), "
          f"mixed={mixed}", flush=True)
    ms = []
    for (name, n, p, chains) in (("c1", 50, 3, 2), ("c1", 7, 2, 1), ("c5_0", 120, 5, 3), ("c5_1", 30, 8, 1)):
        d = synth.generate(name, nsubs=n, ncovar=p)
        ms.append((d, Layout(d["mask"], *d["grid"]), chains))

    def mk():
        out = []
        for k, (d, L, ch) in enumerate(ms):
            zfree = mixed and k % 2 == 1
            out.append(Sampler(L, d["covar"], d["data"], seed=SEED + k, nchains=ch, xtz_mode=2 if zfree else 1,
                               z_store=2 if zfree else 0))
        return out

    solo, grouped = mk(), mk()
    B = Batch(grouped)
    assert not B.one_launch_per_phase
    for S in solo:
        S.sweep(5)
    B.sweep(5)
    for S in solo:
        S.accum_reset(ntypes=1, nbatches=3, batch_len=2)
        S.sweep_accumulate(6, 1)
    for S in grouped:
        S.accum_reset(ntypes=1, nbatches=3, batch_len=2)
    B.sweep_accumulate(6, 1)
    for S, T in zip(solo, grouped):
        S.synchronize()
        T.synchronize()
        assert S.sweep_index() == T.sweep_index() == 11
        assert T.fallback_count() == 0
        for c in range(S.nchains):
            assert np.array_equal(S.get_spat_coef(c), T.get_spat_coef(c))
            assert np.array_equal(S.get_xtz(c), T.get_xtz(c))
            assert np.array_equal(S.get_prec(c), T.get_prec(c))
            a, b = S.accum_get(c), T.accum_get(c)
            for k in ("mean", "var", "prb", "mcse"):
                assert np.array_equal(a[k], b[k]), k
    B.close()


@pytest.mark.parametrize("name,nchains", [("c5_3", 2), ("c1", 1)])
def test_autotune_keeps_the_chain_state(name, nchains):
    """The autotuner times whole sweeps at every lane split from one saved
    state and puts the state back: coefficients, X'Z, Lambda, pair sums and
    the sweep index are unchanged, and the next sweeps equal those of an
    untouched twin pinned to the chosen split."""
    print(f"case autotune {name} chains={nchains}", flush=True)
    d = synth.generate(name)
    L = Layout(d["mask"], *d["grid"])
    S, T = (Sampler(L, d["covar"], d["data"], seed=SEED, nchains=nchains, xtz_mode=2, z_store=2) for _ in range(2))
    for X in (S, T):
        X.sweep(3)
        X.synchronize()

    def state(X):
        return [(X.get_spat_coef(c), X.get_xtz(c), X.get_prec(c), X.get_pair_sums(c)) for c in range(nchains)]

    before, idx = state(S), S.sweep_index()
    g, ms = S.autotune(reps=2)
    assert g in (1, 2, 4, 8) and ms[[1, 2, 4, 8].index(g)] == ms.min()
    assert S.sweep_index() == idx
    for a, b in zip(state(S), before):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    T.lane_split = g
    S.sweep(2)  # sweeps run with the tuned split
    T.sweep(2)
    for a, b in zip(state(S), state(T)):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_batch_rejects_unsupported_members():
    d = synth.generate("c1", nsubs=9, ncovar=2)
    L = Layout(d["mask"], *d["grid"])
    with pytest.raises(Exception):
        Batch([Sampler(L, d["covar"], d["data"], seed=1, xtz_mode=2, z_store=0)])
    A = Sampler(L, d["covar"], d["data"], seed=1, xtz_mode=2, z_store=2)
    A.sweep(1)
    with pytest.raises(Exception):  # sweep indices differ
        Batch([A, Sampler(L, d["covar"], d["data"], seed=2, xtz_mode=2, z_store=2)])


def test_batch_refuses_members_swept_on_their_own():
    """The batch draws every member's Philox words from one sweep index: a
    member swept alone after the batch was formed is refused, not reused."""
    print("case batch after a member swept alone", flush=True)
    d = synth.generate("c1", nsubs=9, ncovar=2)
    L = Layout(d["mask"], *d["grid"])
    A = Sampler(L, d["covar"], d["data"], seed=1, xtz_mode=2, z_store=2)
    Bm = Sampler(L, d["covar"], d["data"], seed=2, xtz_mode=2, z_store=2)
    B = Batch([A, Bm])
    B.sweep(2)
    A.sweep(1)
    with pytest.raises(Exception, match="sweep indices"):
        B.sweep(1)
    Bm.sweep(1)  # back at one index: the batch runs again
    B.sweep(1)
    assert A.sweep_index() == Bm.sweep_index() == 4
    B.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_batch_graph_replays_equal_direct_batch_sweeps(mode):
    """bsg_batch_sweep_graph: groups of batch sweeps replayed from one HIP
    graph (the stream of each covariate count forked from and joined to one
    origin stream; the sweep word advanced by the graph's last node) equal
    direct batch sweeps bit for bit, with direct sweeps and reads between
    replays."""
    print("case batch graph vs direct, 4 analyses, three covariate counts", flush=True)
    ms = members()
    direct, graphed = make(ms), make(ms)
    BD, BG = Batch(direct), Batch(graphed)
    BG.set_launch_mode(mode)
    for step in (10, 3, 8):  # 10: two replays of 4 + 2 direct; 3: direct only; 8: two replays
        BD.sweep(step)
        BG.sweep_graph(step, 4)
        for S, T in zip(direct, graphed):
            S.synchronize()
            T.synchronize()
            assert S.sweep_index() == T.sweep_index()
            for c in range(S.nchains):
                assert np.array_equal(S.get_spat_coef(c), T.get_spat_coef(c))
                assert np.array_equal(S.get_prec(c), T.get_prec(c))
    BD.close()
    BG.close()
