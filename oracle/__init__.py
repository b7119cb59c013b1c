"""TEST INFRASTRUCTURE ONLY: Python bindings of the checker libraries.

  * liboracle.so       -- FP64 restatement of the sweep (bsg_oracle.c)
  * _ref/libbsg_ref.so -- the reference's own CPU-path sources compiled in
                          place (bsglmm/covar.cpp, randgen.cpp, cholesky.cpp)

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
import this package. The product (bsglmm_amd) never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libbsg_ref.so")

_orc = None
_ref = None

pi, pf, pd, pu8, pu32 = (C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_double),
                        C.POINTER(C.c_uint8), C.POINTER(C.c_uint32))


def build():
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Problem(C.Structure):
    _fields_ = [("nrow", C.c_int), ("ncol", C.c_int), ("ndepth", C.c_int),
                ("nsubs", C.c_int), ("ncovar", C.c_int), ("totvox", C.c_int), ("totvoxp", C.c_int),
                ("n_col", C.c_int * 2), ("host_idx", pi), ("host_idx_sc", pi),
                ("vox", pi * 2), ("nbrs", pu8 * 2), ("covar", pf), ("xxprime", pf),
                ("data", pu8), ("wm", pf), ("key", C.c_uint32 * 2), ("chain", C.c_uint32)]


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        sig = {
            "orc_philox4x32_10": (None, [pu32, pu32, pu32]),
            "orc_draw": (None, [pu32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, pu32]),
            "orc_u01": (C.c_double, [C.c_uint32]),
            "orc_Q": (C.c_double, [C.c_double]),
            "orc_Qinv": (C.c_double, [C.c_double]),
            "orc_logQ": (C.c_double, [C.c_double]),
            "orc_latent_draw": (C.c_double, [C.c_double, C.c_int, C.c_double, C.c_double]),
            "orc_layout_counts": (None, [pu8, C.c_int, C.c_int, C.c_int, pi]),
            "orc_layout_fill": (None, [pu8, C.c_int, C.c_int, C.c_int, pi, pi, pi, pi, pu8, pu8, pi]),
            "orc_update_z": (None, [C.POINTER(Problem), pf, C.c_float, C.c_double, C.c_uint32, C.c_int, pd]),
            "orc_xtz": (None, [C.POINTER(Problem), pf, C.c_float, pd]),
            "orc_update_coef": (None, [C.POINTER(Problem), C.c_int, pd, pf, pd, C.c_uint32, pd]),
            "orc_pair_sums": (None, [C.POINTER(Problem), pf, pd]),
            "orc_wm_stat": (None, [C.POINTER(Problem), pf, pd, pd, pd]),
            "orc_wm_draw": (C.c_double, [C.c_double, C.c_double, pu32, C.c_uint32, C.c_uint32]),
            "orc_cpu_wm_stat": (None, [C.POINTER(Problem), pf, pf, pf, C.c_float, C.c_float, pf, pf]),
            "orc_wishart": (None, [C.c_int, pd, C.c_int, pu32, C.c_uint32, C.c_uint32, pd, pd, pd]),
            "orc_gamma": (C.c_double, [C.c_double, pu32, C.c_uint32, C.c_uint32, C.c_uint32]),
            "orc_chol": (C.c_int, [pd, C.c_int]),
            "orc_chol_invert": (None, [C.c_int, pd]),
            "orc_kiss": (C.c_double, [C.POINTER(C.c_ulong)]),
            "orc_runiform_long_n": (C.c_ulong, [C.c_ulong, C.POINTER(C.c_ulong)]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype, f.argtypes = r, a
        _orc = L
    return _orc


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    """The reference's compiled CPU path, or None when oracle/_ref is not built."""
    global _ref
    if _ref is None:
        if not ref_available():
            return None
        L = C.CDLL(REF_SO)
        pul = C.POINTER(C.c_ulong)
        sig = {
            "ref_setup": (C.c_int, [C.c_int] * 9 + [pi, pi, pi, pi, pu8, pu8, pf]),
            "ref_kiss": (C.c_double, [pul]),
            "ref_snorm": (C.c_double, [pul]),
            "ref_rgamma": (C.c_double, [C.c_double, C.c_double, pul]),
            "ref_truncnorm2": (C.c_double, [C.c_double] * 4 + [pul]),
            "ref_runiform_long_n": (C.c_ulong, [C.c_ulong, pul]),
            "ref_rwishart2": (C.c_int, [pd, pd, C.c_int, C.c_int, pul]),
            "ref_cholesky_decomp": (C.c_int, [pd, C.c_int]),
            "ref_cholesky_invert": (None, [C.c_int, pd]),
            "ref_update_z": (None, [pf, pu8, C.c_float, pf, pf, pf, pul]),
            "ref_update_spat_coef": (None, [pf, pf, pf, pf, C.c_float, pf, pul]),
            "ref_update_spat_coef_prec": (None, [pf, pf, pul]),
            "ref_spat_coef_precision": (C.c_double, [pf, pf, C.c_int]),
            "ref_sweep": (None, [pf, pu8, pf, pf, pf, pf, pul]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype, f.argtypes = r, a
        _ref = L
    return _ref


# ---------------------------------------------------------------------------
def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.empty(4, np.uint32)
    orc().orc_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(out, C.c_uint32))
    return out


def layout(mask, nrow, ncol, ndepth):
    m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8).reshape(-1))
    cnt = np.empty(4, np.int32)
    orc().orc_layout_counts(_p(m, C.c_uint8), nrow, ncol, ndepth, _p(cnt, C.c_int))
    V, VP, n0, n1 = (int(x) for x in cnt)
    nsize = m.size
    out = dict(nrow=nrow, ncol=ncol, ndepth=ndepth, totvox=V, totvoxp=VP, n0=n0, n1=n1,
               host_idx=np.empty(nsize, np.int32), host_idx_sc=np.empty(nsize, np.int32),
               vox0=np.empty(max(n0, 1), np.int32), vox1=np.empty(max(n1, 1), np.int32),
               nbrs0=np.empty(max(n0, 1), np.uint8), nbrs1=np.empty(max(n1, 1), np.uint8),
               nbr6=np.empty(max(V, 1) * 6, np.int32))
    orc().orc_layout_fill(_p(m, C.c_uint8), nrow, ncol, ndepth, _p(out["host_idx"], C.c_int),
                          _p(out["host_idx_sc"], C.c_int), _p(out["vox0"], C.c_int),
                          _p(out["vox1"], C.c_int), _p(out["nbrs0"], C.c_uint8),
                          _p(out["nbrs1"], C.c_uint8), _p(out["nbr6"], C.c_int))
    for k, n in (("vox0", n0), ("vox1", n1), ("nbrs0", n0), ("nbrs1", n1)):
        out[k] = out[k][:n]
    out["nbr6"] = out["nbr6"][:V * 6]
    return out


def xtx_reference(covar):
    """X'X accumulated in float32, subject-outer (read_data.cpp:321-327)."""
    P, N = covar.shape
    acc = np.zeros((P, P), np.float32)
    for i in range(N):
        x = covar[:, i].astype(np.float32)
        acc += np.outer(x, x).astype(np.float32)
    return acc


class Oracle:
    """One problem for the FP64 oracle, in the reference's layouts."""

    def __init__(self, lay, covar, data, wm=None, seed=0, chain=0):
        self.lay = lay
        self.covar = np.ascontiguousarray(covar, dtype=np.float32)
        self.P, self.N = self.covar.shape
        self.V, self.VP = lay["totvox"], lay["totvoxp"]
        self.data = np.ascontiguousarray(data, dtype=np.uint8)
        self.wm = np.ascontiguousarray(np.ones(self.V, np.float32) if wm is None else wm, dtype=np.float32)
        self.xtx = xtx_reference(self.covar)
        self.key = np.array([seed & 0xffffffff, (seed >> 32) & 0xffffffff], np.uint32)
        pb = Problem()
        pb.nrow, pb.ncol, pb.ndepth = lay["nrow"], lay["ncol"], lay["ndepth"]
        pb.nsubs, pb.ncovar, pb.totvox, pb.totvoxp = self.N, self.P, self.V, self.VP
        pb.n_col[0], pb.n_col[1] = lay["n0"], lay["n1"]
        pb.host_idx = _p(lay["host_idx"], C.c_int)
        pb.host_idx_sc = _p(lay["host_idx_sc"], C.c_int)
        pb.vox[0], pb.vox[1] = _p(lay["vox0"], C.c_int), _p(lay["vox1"], C.c_int)
        pb.nbrs[0], pb.nbrs[1] = _p(lay["nbrs0"], C.c_uint8), _p(lay["nbrs1"], C.c_uint8)
        pb.covar, pb.xxprime = _p(self.covar, C.c_float), _p(self.xtx, C.c_float)
        pb.data, pb.wm = _p(self.data, C.c_uint8), _p(self.wm, C.c_float)
        pb.key[0], pb.key[1] = int(self.key[0]), int(self.key[1])
        pb.chain = chain
        self.pb = pb

    def set_chain(self, chain):
        self.pb.chain = chain

    def update_z(self, spat_coef, sweep, b=0.0, eps=1e-6, method=0):
        """method 0: inversion for every draw (k_latent); 1: Box-Muller candidate
        with an exact fallback (k_latent_vl, the form without Z)."""
        sc = np.ascontiguousarray(spat_coef, dtype=np.float32)
        z = np.empty((self.N, self.V), np.float64)
        orc().orc_update_z(C.byref(self.pb), _p(sc, C.c_float), b, eps, sweep, int(method), _p(z, C.c_double))
        return z

    def xtz(self, z, b=0.0):
        z = np.ascontiguousarray(z, dtype=np.float32)
        m = np.empty((self.P, self.V), np.float64)
        orc().orc_xtz(C.byref(self.pb), _p(z, C.c_float), b, _p(m, C.c_double))
        return m

    def update_coef(self, colour, m, prec, spat_coef, sweep, z_inj=None):
        m = np.ascontiguousarray(m, dtype=np.float64)
        prec = np.ascontiguousarray(prec, dtype=np.float32)
        sc = np.ascontiguousarray(spat_coef, dtype=np.float64).copy()
        zp = None
        if z_inj is not None:
            z_inj = np.ascontiguousarray(z_inj, dtype=np.float64)
            zp = _p(z_inj, C.c_double)
        orc().orc_update_coef(C.byref(self.pb), colour, _p(m, C.c_double), _p(prec, C.c_float),
                              _p(sc, C.c_double), sweep, zp)
        return sc

    def wm_stat(self, spat_coef, z):
        """(S, W) of the WM coefficient draw from z[N][TOTVOX] (covarGPU.cu:588-683)."""
        sc = np.ascontiguousarray(spat_coef, dtype=np.float32)
        z = np.ascontiguousarray(z, dtype=np.float64)
        S, W = C.c_double(), C.c_double()
        orc().orc_wm_stat(C.byref(self.pb), _p(sc, C.c_float), _p(z, C.c_double), C.byref(S), C.byref(W))
        return S.value, W.value

    def cpu_wm_stat(self, spat_coef, z, prior_mean=0.0, prior_prec=1e-6):
        """(mean, var) of b from the reference CPU path's statistic (updateBeta,
        mcmc.cpp:1099-1136: float Kahan sums in raster order) for float z[N][TOTVOX]."""
        sc = np.ascontiguousarray(spat_coef, dtype=np.float32)
        z = np.ascontiguousarray(z, dtype=np.float32)
        m, v = C.c_float(), C.c_float()
        orc().orc_cpu_wm_stat(C.byref(self.pb), _p(sc, C.c_float), _p(z, C.c_float), _p(self.wm, C.c_float),
                              prior_mean, prior_prec, C.byref(m), C.byref(v))
        return m.value, v.value

    def wm_draw(self, S, W, sweep):
        """b ~ N(S/W, 1/W) from the Philox normal of (0, sweep, BWM, chain) (covarGPU.cu:711-717)."""
        k = (C.c_uint32 * 2)(int(self.key[0]), int(self.key[1]))
        return orc().orc_wm_draw(C.c_double(S), C.c_double(W), k, sweep, int(self.pb.chain))

    def pair_sums(self, spat_coef):
        sc = np.ascontiguousarray(spat_coef, dtype=np.float32)
        t = np.empty((self.P, self.P), np.float64)
        orc().orc_pair_sums(C.byref(self.pb), _p(sc, C.c_float), _p(t, C.c_double))
        return t

    def wishart(self, t, df, sweep, chi2=None, norms=None):
        return wishart(self.P, t, df, self.key, sweep, int(self.pb.chain), chi2, norms)


def wishart(p, t, df, key, sweep, chain=0, chi2=None, norms=None):
    t = np.ascontiguousarray(t, dtype=np.float64)
    key = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.empty((p, p), np.float64)
    cp = _p(np.ascontiguousarray(chi2, dtype=np.float64), C.c_double) if chi2 is not None else None
    npp = _p(np.ascontiguousarray(norms, dtype=np.float64), C.c_double) if norms is not None else None
    orc().orc_wishart(p, _p(t, C.c_double), df, _p(key, C.c_uint32), sweep, chain, cp, npp,
                      _p(out, C.c_double))
    return out


class RefCPU:
    """The reference CPU path (oracle/_ref) set up for one problem."""

    def __init__(self, lay, covar, xtx):
        self.L = ref()
        if self.L is None:
            raise RuntimeError("oracle/_ref is not built")
        self.lay = lay
        self._keep = [lay, np.ascontiguousarray(xtx, dtype=np.float32)]
        P, N = covar.shape
        rc = self.L.ref_setup(lay["nrow"], lay["ncol"], lay["ndepth"], N, P, lay["totvox"], lay["totvoxp"],
                              lay["n0"], lay["n1"], _p(lay["host_idx"], C.c_int),
                              _p(lay["host_idx_sc"], C.c_int), _p(lay["vox0"], C.c_int),
                              _p(lay["vox1"], C.c_int), _p(lay["nbrs0"], C.c_uint8),
                              _p(lay["nbrs1"], C.c_uint8), _p(self._keep[1], C.c_float))
        assert rc == 0


def seed_array(triple):
    return (C.c_ulong * 3)(*[int(x) for x in triple])
