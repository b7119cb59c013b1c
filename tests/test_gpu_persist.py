"""Persistent sweep (k_persist): the sweeps of one bsg_sweep call in one
launch, resident blocks taking latent tiles, colour updates, Bartlett factors
and CAR-precision draws from a queue, values passed between blocks as tagged
words. It must reproduce the overlapped and the serial sweeps bit for bit
(the same Philox counters, tiles, pieces and sums; mcmc.cpp:311-394 order of
updates): coefficients, X'Z, Lambda, pair sums and the sweep index, across
call boundaries, at any queue lags, through the FP64 refactorisation, and
over long runs (where a missing wait would show as a race)."""
import numpy as np
import pytest

import oracle
from bsglmm_amd import Layout, Sampler, lib, synth

pytestmark = pytest.mark.gpu

SEED = 0x5EED9E3779B97F4A


def spat_coef_from_ranks(beta, lay):
    hi, hs = lay["host_idx"], lay["host_idx_sc"]
    ids = np.flatnonzero(hi >= 0)
    sc = np.zeros((beta.shape[0], lay["totvoxp"]), np.float32)
    sc[:, hs[ids]] = beta[:, hi[ids]]
    return sc


def make(name, nsubs, ncovar, forms, g=0, **kw):
    """Samplers of one problem and start state; forms: 'persistent',
    'overlap' or 'serial' each."""
    args = {} if name in ("c2", "c3") else dict(nsubs=nsubs, ncovar=ncovar)
    d = synth.generate(name, **args)
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    L = Layout(d["mask"], nx, ny, nz)
    rng = np.random.default_rng(5)
    sc = spat_coef_from_ranks(d["beta_true"] + 0.05 * rng.standard_normal(d["beta_true"].shape)
                              .astype(np.float32), lay)
    out = []
    for f in forms:
        S = Sampler(L, d["covar"], d["data"], seed=SEED, xtz_mode=2, z_store=2, **kw)
        if g:
            S.lane_split = g
        S.overlap = f != "serial"
        S.persistent = f == "persistent"
        assert S.persistent == (f == "persistent"), f
        S.set_spat_coef(sc)
        S.set_sweep_index(11)
        out.append(S)
    return out


def assert_same_state(A, B):
    A.synchronize()
    B.synchronize()
    assert A.sweep_index() == B.sweep_index()
    assert np.array_equal(A.get_spat_coef(), B.get_spat_coef())
    assert np.array_equal(A.get_xtz(), B.get_xtz())
    assert np.array_equal(A.get_prec(), B.get_prec())
    assert np.array_equal(A.get_pair_sums(), B.get_pair_sums())


# (N, P, lane split) as the overlap tests: P = 1..6, leftover tiles in pieces
# (N >= 57) or none (N = 13), C2 at full size
@pytest.mark.parametrize("name,nsubs,ncovar,g", [("c1", 300, 4, 0), ("c1", 200, 6, 0), ("c1", 57, 5, 4),
                                                 ("c1", 13, 3, 4), ("c1", 90, 2, 0), ("c1", 150, 1, 0),
                                                 ("c2", None, None, 0)])
def test_persistent_sweep_equals_overlapped_and_serial(name, nsubs, ncovar, g):
    print(f"case persistent vs overlapped vs serial {name} N={nsubs} P={ncovar} G={g}", flush=True)
    A, B, C = make(name, nsubs, ncovar, ("persistent", "overlap", "serial"), g)
    A.sweep(1)  # calls of 1, 2 and 4 sweeps: each starts from what the one before left
    A.sweep(2)
    A.sweep(4)
    B.sweep(7)
    C.sweep(7)
    assert_same_state(A, B)
    assert_same_state(A, C)
    assert A.fallback_count() == C.fallback_count()


def test_persistent_lags_change_no_result():
    """The queue lags move items in the queue, never their inputs: three
    settings (short, the default, long) give the same chain."""
    print("case persistent lags c1 N=300 P=4", flush=True)
    A, B, C = make("c1", 300, 4, ("persistent", "persistent", "persistent"))
    A.set_persistent_lags(0.02, 0.02, 0.02)
    C.set_persistent_lags(4.0, 1.0, 1.0)
    for S in (A, B, C):
        S.sweep(9)
    assert_same_state(A, B)
    assert_same_state(A, C)


def test_persistent_fp64_refactor_every_voxel():
    """pivot_min = 1e6 sends every voxel through the FP64 refactorisation
    inside the persistent kernel's updates."""
    print("case persistent FP64 fallback c1 N=150 P=4", flush=True)
    A, B = make("c1", 150, 4, ("persistent", "serial"), pivot_min=1e6)
    A.sweep(3)
    B.sweep(3)
    assert_same_state(A, B)
    assert A.fallback_count() == B.fallback_count() > 0


def test_persistent_long_run_c2():
    """200 sweeps of the full C2 problem in one launch: any item that ran
    before an input it needs would break the equality."""
    print("case persistent long run c2, 200 sweeps", flush=True)
    A, B = make("c2", None, None, ("persistent", "overlap"))
    A.sweep(200)
    B.sweep(200)
    assert_same_state(A, B)


def test_persistent_falls_back_where_unsupported():
    """Save sweeps and graphs run overlapped; several chains and P = 7 never
    run persistent. The chains are unchanged."""
    print("case persistent fallbacks c1 N=300 P=4", flush=True)
    A, B = make("c1", 300, 4, ("persistent", "overlap"))
    for S in (A, B):
        S.accum_reset(ntypes=1, nbatches=2, batch_len=2)
        S.sweep_accumulate(4, save_every=1)
    A.sweep_graph(4, per_graph=2)
    B.sweep(4)
    A.sweep(3)
    B.sweep(3)
    assert_same_state(A, B)
    ra, rb = A.accum_get(), B.accum_get()
    for k in ra:
        assert np.array_equal(np.asarray(ra[k]), np.asarray(rb[k]), equal_nan=True), k
    d = synth.generate("c1", nsubs=60, ncovar=7)
    L = Layout(d["mask"], *d["grid"])
    S = Sampler(L, d["covar"], d["data"], seed=1, xtz_mode=2, z_store=2)
    S.overlap = True
    S.persistent = True
    assert not S.persistent
    d = synth.generate("c1", nsubs=60, ncovar=3)
    L = Layout(d["mask"], *d["grid"])
    S = Sampler(L, d["covar"], d["data"], seed=1, nchains=2, xtz_mode=2, z_store=2)
    S.persistent = True
    assert S.overlap and not S.persistent


def test_persistent_bounded_wait_reports_and_recovers():
    """Every wait of the persistent kernel is bounded. Forced from the host
    (bsg_test_overlap_wait): the draw writes a tag no reader accepts, the
    updates of the second sweep give up after 8 polls, so does every later
    wait, the launch ends, and bsg_synchronize reports it once; with the hook
    cleared the next launch (a new epoch of tags) sweeps on."""
    print("case persistent bounded wait forced from the host", flush=True)
    (A,) = make("c1", 300, 4, ("persistent",))
    assert lib().bsg_test_overlap_wait(A._h, 8, 2) == 0
    A.sweep(3)
    with pytest.raises(Exception, match="past its bound"):
        A.synchronize()
    assert lib().bsg_test_overlap_wait(A._h, 0, 0) == 0
    (B,) = make("c1", 300, 4, ("serial",))
    A.set_spat_coef(B.get_spat_coef())
    A.set_prec(B.get_prec())
    A.set_sweep_index(B.sweep_index())
    A.sweep(2)
    B.sweep(2)
    assert_same_state(A, B)
