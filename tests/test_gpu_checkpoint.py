"""Checkpoint and resume (the reference's save_iter / restart_iter,
mcmc.cpp:1265-1336, made exact): k sweeps, a checkpoint, a new context read
from it and n - k more sweeps must equal n sweeps of one context bit for bit:
coefficients, CAR precision, pair sums, WM coefficient, X'Z, sweep index and
the posterior accumulators. The restored context may be created with another
seed: the saved Philox key replaces it."""
import os

import numpy as np
import pytest

from bsglmm_amd import BsgError, Layout, Sampler, synth

pytestmark = pytest.mark.gpu


def state(S):
    out = []
    for c in range(S.nchains):
        a = S.accum_get(c)
        out.append((S.get_spat_coef(c), S.get_prec(c), S.get_pair_sums(c), S.get_wm_coef(c), S.get_xtz(c),
                    a["ndraws"], a["mean"], a["var"], a["prb"], a["mcse"], a["bwm"]))
    return S.sweep_index(), out


def assert_same(a, b):
    assert a[0] == b[0]
    for x, y in zip(a[1], b[1]):
        for u, v in zip(x, y):
            assert np.array_equal(np.asarray(u), np.asarray(v))


@pytest.mark.parametrize("nsubs,ncovar,nchains,z_store,xtz_mode,wm,g",
                         [(150, 4, 1, 2, 2, False, 0),   # overlapped, colour tiles and pieces
                          (40, 3, 3, 0, 0, True, 0),     # serial, Z stored, WM coefficient, 3 chains
                          (300, 6, 2, 2, 2, False, 8)])  # Z-free at eight lanes per voxel
def test_resume_equals_uninterrupted(tmp_path, nsubs, ncovar, nchains, z_store, xtz_mode, wm, g):
    print(f"case checkpoint c1 N={nsubs} P={ncovar} chains={nchains} z_store={z_store} wm={wm} G={g}", flush=True)
    d = synth.generate("c1", nsubs=nsubs, ncovar=ncovar)
    L = Layout(d["mask"], *d["grid"])
    w = np.random.default_rng(2).uniform(0.2, 1.0, L.totvox).astype(np.float32) if wm else None

    def mk(seed=0xC4EC):
        S = Sampler(L, d["covar"], d["data"], wm=w, seed=seed, nchains=nchains, xtz_mode=xtz_mode, z_store=z_store)
        if g:
            S.lane_split = g
        return S

    ref, A = mk(), mk()
    for X in (ref, A):
        X.sweep(4)
        X.accum_reset(ntypes=2, nbatches=4, batch_len=3)
    ref.sweep_accumulate(17, 1)
    A.sweep_accumulate(7, 1)
    path = os.path.join(tmp_path, "chain.bsg")
    A.checkpoint(path)
    A.close()
    B = Sampler(L, d["covar"], d["data"], wm=w, seed=0x0DD5EED, nchains=nchains, xtz_mode=xtz_mode, z_store=z_store)
    B.restore(path)
    assert B.sweep_index() == 11
    assert B.lane_split == (g or B.lane_split)
    B.sweep_accumulate(10, 1)
    ref.synchronize()
    B.synchronize()
    assert_same(state(ref), state(B))
    # and plain sweeps after that
    ref.sweep(3)
    B.sweep(3)
    assert np.array_equal(ref.get_spat_coef(0), B.get_spat_coef(0))


def test_checkpoint_refuses_other_problems_and_damaged_files(tmp_path):
    print("case checkpoint errors", flush=True)
    d = synth.generate("c1", nsubs=30, ncovar=3)
    L = Layout(d["mask"], *d["grid"])
    A = Sampler(L, d["covar"], d["data"], seed=5, xtz_mode=2, z_store=2)
    A.sweep(2)
    path = os.path.join(tmp_path, "a.bsg")
    A.checkpoint(path)
    d4 = synth.generate("c1", nsubs=30, ncovar=4)
    with pytest.raises(BsgError, match="another problem"):
        Sampler(L, d4["covar"], d4["data"], seed=5, xtz_mode=2, z_store=2).restore(path)
    with pytest.raises(BsgError, match="another problem"):
        Sampler(L, d["covar"], d["data"], seed=5, nchains=2, xtz_mode=2, z_store=2).restore(path)
    with pytest.raises(BsgError, match="options"):
        Sampler(L, d["covar"], d["data"], seed=5, xtz_mode=2, z_store=2, df_offset=3).restore(path)
    raw = open(path, "rb").read()
    bad = bytearray(raw)
    bad[len(raw) // 2] ^= 0x40
    open(path, "wb").write(bytes(bad))
    B = Sampler(L, d["covar"], d["data"], seed=5, xtz_mode=2, z_store=2)
    with pytest.raises(BsgError, match="checksum"):
        B.restore(path)
    open(path, "wb").write(raw[:-100])
    with pytest.raises(BsgError, match="truncated|checksum"):
        B.restore(path)
    with pytest.raises(BsgError, match="cannot open"):
        B.restore(os.path.join(tmp_path, "missing.bsg"))
    # a failed read leaves the context usable and unchanged
    assert B.sweep_index() == 0
    B.sweep(1)
    B.synchronize()
