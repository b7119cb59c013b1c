"""bsg_coef_means_noalloc, the allocation-free form of bsg_coef_means, called
through the C ABI: its plane means must equal bsg_coef_means's bit for bit on
two calls in a row, the second on a smaller problem (fewer voxels, subjects,
covariates and chains), so that nothing left in the shared scratch array by
the first call reaches the second. It runs before bsg_coef_means on each
context, i.e. before that entry point has allocated anything."""
import ctypes as C

import numpy as np
import pytest

from bsglmm_amd import Layout, Sampler, lib, synth

pytestmark = pytest.mark.gpu


def means(S, chain, noalloc):
    out = np.empty(S.P, np.float64)
    f = lib().bsg_coef_means_noalloc if noalloc else lib().bsg_coef_means
    rc = f(S._h, chain, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0, lib().bsg_last_error().decode()
    return out


@pytest.mark.parametrize("order", ["large_then_small", "small_then_large"])
def test_noalloc_coef_means_equal_allocating_entry_point(order):
    print(f"case coef_means_noalloc {order}: c2 N=250 P=8 x2 chains, c1 N=40 P=3", flush=True)
    big = synth.generate("c2", ncovar=8)
    small = synth.generate("c1", nsubs=40, ncovar=3)
    cases = [(big, 2), (small, 1)]
    if order == "small_then_large":
        cases.reverse()
    for d, chains in cases:
        L = Layout(d["mask"], *d["grid"])
        S = Sampler(L, d["covar"], d["data"], seed=0xA110C, nchains=chains, xtz_mode=2, z_store=2)
        S.sweep(3)
        S.synchronize()
        chain = chains - 1
        got = means(S, chain, True)
        ref = means(S, chain, False)
        assert np.array_equal(got, ref), (got, ref)
        # and the same as the FP64 mean of the coefficients read back
        sc = S.get_spat_coef(chain).astype(np.float64)
        e = L.export()
        hs = e["host_idx_sc"][e["host_idx"] >= 0]
        assert np.allclose(got, sc[:, hs].mean(1), rtol=1e-12, atol=1e-12)
        S.close()


def test_noalloc_rejects_bad_arguments():
    d = synth.generate("c1", nsubs=20, ncovar=2)
    L = Layout(d["mask"], *d["grid"])
    S = Sampler(L, d["covar"], d["data"], seed=1, xtz_mode=2, z_store=2)
    out = np.empty(2, np.float64)
    assert lib().bsg_coef_means_noalloc(S._h, 1, out.ctypes.data_as(C.POINTER(C.c_double))) != 0
    assert lib().bsg_coef_means_noalloc(S._h, 0, None) != 0
    assert lib().bsg_coef_means_noalloc(None, 0, out.ctypes.data_as(C.POINTER(C.c_double))) != 0
