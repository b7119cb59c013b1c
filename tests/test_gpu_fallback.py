"""GPU parity of the coefficient update's FP64 refactorisation and of its
non-positive-definite branch (coef_fp64 / ul_draw in bsg_kernels.hip).

A voxel whose FP32 Cholesky pivot falls below pivot_min is redone in FP64 from
the same inputs. Raising pivot_min above every pivot sends every voxel there,
so the FP64 branch is checked against the FP64 oracle in each kernel that
carries it: the serial colour kernels (k_coef), the overlapped sweep with one
chain (k_overlap, colour tiles and pieces) and with three chains, and a batch
(k_coef_multi). If FP64 fails too, the precision is not positive definite:
the voxel keeps its coefficients and bsg_synchronize reports the count once,
where the reference propagates NaNs silently (covarGPU.cu:374).

Tolerance of the FP64 branch: 1e-6 relative (floor 1). Its arithmetic is
FP64; the error left is that of its FP32 inputs (the neighbour sum, the
Box-Muller normals on the hardware log/sin/cos), against 1e-5 for the FP32
path in the other parity tests.
"""
import numpy as np
import pytest

import oracle
from bsglmm_amd import Batch, BsgError, Layout, Sampler, synth

pytestmark = pytest.mark.gpu
SEED = 0x0F64FA11BAC4
FP64_TOL = 1e-6


def spat_coef_from_ranks(beta, lay):
    hi, hs = lay["host_idx"], lay["host_idx_sc"]
    ids = np.flatnonzero(hi >= 0)
    sc = np.zeros((beta.shape[0], lay["totvoxp"]), np.float32)
    sc[:, hs[ids]] = beta[:, hi[ids]]
    return sc


def rel_err(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


def setup(nsubs, ncovar, name="c1"):
    d = synth.generate(name, nsubs=nsubs, ncovar=ncovar)
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    L = Layout(d["mask"], nx, ny, nz)
    return d, lay, L


def start_state(S, d, lay, nchains, sweep, rng):
    P = S.P
    prec0 = (np.eye(P) * 1.5 + 0.2).astype(np.float32)
    S.set_sweep_index(sweep)
    scs = []
    for c in range(nchains):
        beta = d["beta_true"] + 0.1 * rng.standard_normal(d["beta_true"].shape).astype(np.float32)
        scs.append(spat_coef_from_ranks(beta, lay))
        S.set_spat_coef(scs[c], chain=c)
        S.set_prec(prec0, chain=c)
    return prec0, scs


def check_against_oracle(S, O, scs, prec0, sweep, nchains, tol):
    worst = 0.0
    for c in range(nchains):
        O.set_chain(c)
        m = S.get_xtz(c).astype(np.float64)
        ref = O.update_coef(0, m, prec0, scs[c].astype(np.float64), sweep)
        ref = O.update_coef(1, m, prec0, ref, sweep)
        e = rel_err(S.get_spat_coef(c), ref)
        worst = max(worst, e)
        assert e < tol, (c, e)
    return worst


# P <= 6 redoes a voxel in registers (coef_fp64), P = 7, 8 in an LDS area of
# the wave (coef_fp64_redo); both the same FP64 arithmetic
@pytest.mark.parametrize("form,nchains,ncovar", [("serial_z_stored", 1, 4), ("serial_z_free", 1, 4), ("overlap", 1, 4),
                                                 ("overlap", 3, 4), ("serial_z_free", 1, 8), ("overlap", 1, 8),
                                                 ("overlap", 2, 7), ("serial_z_stored", 1, 7)])
def test_fp64_refactor_every_voxel(form, nchains, ncovar):
    """pivot_min = 1e6: every voxel of both colours takes the FP64 branch."""
    print(f"case FP64 fallback form={form} chains={nchains} (c1 N=150 P={ncovar})", flush=True)
    d, lay, L = setup(150, ncovar)
    zs = 0 if form == "serial_z_stored" else 2
    xm = 0 if form == "serial_z_stored" else 2
    S = Sampler(L, d["covar"], d["data"], seed=SEED, nchains=nchains, xtz_mode=xm, z_store=zs, pivot_min=1e6)
    T = Sampler(L, d["covar"], d["data"], seed=SEED, nchains=nchains, xtz_mode=xm, z_store=zs)
    for X in (S, T):
        X.overlap = form == "overlap"
    assert S.overlap == (form == "overlap")
    O = oracle.Oracle(lay, d["covar"], d["data"], seed=SEED)
    sweep = 4
    prec0, scs = start_state(S, d, lay, nchains, sweep, np.random.default_rng(3))
    start_state(T, d, lay, nchains, sweep, np.random.default_rng(3))
    if form == "serial_z_stored":
        for X in (S, T):  # the reference's step order (mcmc.cpp:311-394)
            X.update_latent()
            X.compute_xtz()
            X.update_coef(0)
            X.update_coef(1)
    else:
        S.sweep(1)
        T.sweep(1)
    S.synchronize()
    T.synchronize()
    assert S.fallback_count() == nchains * L.totvox
    assert T.fallback_count() == 0
    e64 = check_against_oracle(S, O, scs, prec0, sweep, nchains, FP64_TOL)
    e32 = check_against_oracle(T, O, scs, prec0, sweep, nchains, 1e-5)
    print(f"max relative error vs FP64 oracle: FP64 branch {e64:.2e}, FP32 path {e32:.2e}", flush=True)


def test_fp64_refactor_in_a_batch():
    """k_coef_multi: every voxel of every member refactored in FP64 inside a
    batch; each member against the oracle, and the batch against the members
    swept alone."""
    print("case FP64 fallback in a batch (c1 N=150 P=4 x2 chains, c1 N=60 P=6, c1 N=40 P=8)", flush=True)
    rows = []
    for k, (n, p, ch) in enumerate(((150, 4, 2), (60, 6, 1), (40, 8, 1))):
        d, lay, L = setup(n, p)
        mk = lambda: Sampler(L, d["covar"], d["data"], seed=SEED + k, nchains=ch, xtz_mode=2, z_store=2,
                             pivot_min=1e6)
        rows.append((d, lay, L, ch, mk(), mk()))
    sweep = 2
    states = []
    for d, lay, L, ch, A, B in rows:
        prec0, scs = start_state(A, d, lay, ch, sweep, np.random.default_rng(9))
        start_state(B, d, lay, ch, sweep, np.random.default_rng(9))
        states.append((prec0, scs))
    bat = Batch([r[4] for r in rows])
    bat.sweep(1)
    for r in rows:
        r[5].sweep(1)
    for k, ((d, lay, L, ch, A, B), (prec0, scs)) in enumerate(zip(rows, states)):
        A.synchronize()
        B.synchronize()
        assert A.fallback_count() == ch * L.totvox
        O = oracle.Oracle(lay, d["covar"], d["data"], seed=SEED + k)
        check_against_oracle(A, O, scs, prec0, sweep, ch, FP64_TOL)
        for c in range(ch):
            assert np.array_equal(A.get_spat_coef(c), B.get_spat_coef(c))
            assert np.array_equal(A.get_prec(c), B.get_prec(c))
    bat.close()


@pytest.mark.parametrize("overlap", [False, True])
def test_not_positive_definite_keeps_coefficients(overlap):
    """Lambda = -1000 I: X'X + n_j Lambda is not positive definite for every
    voxel with a neighbour, in FP32 and in FP64. Those voxels keep their
    coefficients, bsg_synchronize reports their count once (BSG_ERR_NUM),
    and the next synchronisation is clean."""
    print(f"case non-SPD precision, overlap={overlap} (c1 N=150 P=4)", flush=True)
    d, lay, L = setup(150, 4)
    S = Sampler(L, d["covar"], d["data"], seed=SEED, xtz_mode=2, z_store=2)
    S.overlap = overlap
    assert S.overlap == overlap
    P = S.P
    sc0 = spat_coef_from_ranks(d["beta_true"], lay)
    S.set_spat_coef(sc0)
    bad = (-1000.0 * np.eye(P)).astype(np.float32)
    S.set_prec(bad)
    S.sweep(1)
    # which voxels cannot be factored: by their neighbour count n_j
    xtx = oracle.xtx_reference(d["covar"]).astype(np.float64)

    def spd(n):
        try:
            np.linalg.cholesky(xtx + n * bad.astype(np.float64))
            return True
        except np.linalg.LinAlgError:
            return False

    hi, hs = lay["host_idx"], lay["host_idx_sc"]
    fails = np.zeros(lay["totvoxp"], bool)
    for c in (0, 1):
        for vox, n in zip(lay[f"vox{c}"], lay[f"nbrs{c}"]):
            fails[hs[vox]] = not spd(int(n))
    nfail = int(fails.sum())
    assert nfail > 0.9 * L.totvox
    with pytest.raises(BsgError, match=f"{nfail} voxel"):
        S.synchronize()
    S.synchronize()  # reported once
    after = S.get_spat_coef()
    assert np.array_equal(after[:, fails], sc0[:, fails])
    ok = np.zeros(lay["totvoxp"], bool)
    ok[hs[np.flatnonzero(hi >= 0)]] = True
    ok &= ~fails
    if ok.any():  # voxels without neighbours update as usual
        assert not np.array_equal(after[:, ok], sc0[:, ok])
    assert np.all(np.isfinite(after))
    assert S.fallback_count() == nfail
