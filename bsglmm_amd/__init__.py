"""bsglmm_amd -- MI355X-native Gibbs sweep of BSGLMM behind a C ABI.

The compute path is libbsglmm.so (hand-written HIP kernels for gfx950, see
include/bsglmm.h). This module is a thin ctypes binding over that ABI, shaped
like the reference's update functions (bsglmm/covarGPU.cu):

    Layout(mask, nrow, ncol, ndepth)          main.cu:189-394 layout
    Sampler.update_latent()                   updateZGPU          covarGPU.cu:1625
    Sampler.update_wm_coef()                  updateBetaGPU       covarGPU.cu:686
    Sampler.compute_xtz() / update_coef(c)    updateSpatCoefGPU   covarGPU.cu:535
    Sampler.update_car_prec()                 updateSpatCoefPrecGPU covarGPU.cu:1063
    Sampler.sweep(n)                          mcmc.cpp:311-394 loop body

There is no CPU fallback: if the library is missing, importing the binding
raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# BSGLMM_LIB selects another build of the library (for example one built with
# host sanitizers); the default is the in-tree build
LIB_PATH = os.environ.get("BSGLMM_LIB") or os.path.join(HERE, "libbsglmm.so")

_lib = None


class BsgError(RuntimeError):
    pass


class _Opts(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("nchains", C.c_int), ("xtz_mode", C.c_int),
                ("df_offset", C.c_int), ("trunc_eps", C.c_float), ("pivot_min", C.c_float),
                ("z_store", C.c_int), ("latent_method", C.c_int)]


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def lib():
    """Load libbsglmm.so (built in-tree by bsglmm_amd.build). Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BsgError("libbsglmm.so is not built: run `python -m bsglmm_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    pi, pf, pd, pu8 = C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_uint8)
    sig = {
        "bsg_last_error": (C.c_char_p, []),
        "bsg_layout_create": (C.c_int, [pu8, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
        "bsg_layout_destroy": (None, [vp]),
        "bsg_layout_counts": (C.c_int, [vp, pi]),
        "bsg_layout_export": (C.c_int, [vp, pi, pi, pi, pi, pu8, pu8, pi]),
        "bsg_opts_default": (None, [C.POINTER(_Opts)]),
        "bsg_create": (C.c_int, [vp, C.c_int, C.c_int, pf, pu8, pf, C.POINTER(_Opts), C.POINTER(vp)]),
        "bsg_destroy": (C.c_int, [vp]),
        "bsg_sweep": (C.c_int, [vp, C.c_int]),
        "bsg_sweep_graph": (C.c_int, [vp, C.c_int, C.c_int]),
        "bsg_sweep_graph_prepare": (C.c_int, [vp, C.c_int]),
        "bsg_update_latent": (C.c_int, [vp]),
        "bsg_compute_xtz": (C.c_int, [vp]),
        "bsg_update_wm_coef": (C.c_int, [vp]),
        "bsg_get_wm_coef": (C.c_int, [vp, C.c_int, pf]),
        "bsg_set_wm_coef": (C.c_int, [vp, C.c_int, C.c_float]),
        "bsg_update_coef": (C.c_int, [vp, C.c_int]),
        "bsg_update_car_prec": (C.c_int, [vp]),
        "bsg_synchronize": (C.c_int, [vp]),
        "bsg_sweep_profiled": (C.c_int, [vp, C.c_int, pd]),
        "bsg_sweep_timed": (C.c_int, [vp, C.c_int, pd, pd]),
        "bsg_accum_reset": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "bsg_accum_add": (C.c_int, [vp]),
        "bsg_accum_shape": (C.c_int, [vp, pi, pi, pi]),
        "bsg_sweep_accumulate": (C.c_int, [vp, C.c_int, C.c_int]),
        "bsg_accum_get": (C.c_int, [vp, C.c_int, pi, pd, pd, pd, pd, pd]),
        "bsg_get_spat_coef": (C.c_int, [vp, C.c_int, pf]),
        "bsg_set_spat_coef": (C.c_int, [vp, C.c_int, pf]),
        "bsg_get_z": (C.c_int, [vp, C.c_int, pf]),
        "bsg_get_xtz": (C.c_int, [vp, C.c_int, pf]),
        "bsg_set_xtz": (C.c_int, [vp, C.c_int, pf]),
        "bsg_get_prec": (C.c_int, [vp, C.c_int, pf]),
        "bsg_set_prec": (C.c_int, [vp, C.c_int, pf]),
        "bsg_get_pair_sums": (C.c_int, [vp, C.c_int, pd]),
        "bsg_get_sweep_index": (C.c_int, [vp, C.c_int, C.POINTER(C.c_uint32)]),
        "bsg_set_sweep_index": (C.c_int, [vp, C.c_uint32]),
        "bsg_get_fallback_count": (C.c_int, [vp, C.POINTER(C.c_uint32)]),
        "bsg_checkpoint_write": (C.c_int, [vp, C.c_char_p]),
        "bsg_checkpoint_read": (C.c_int, [vp, C.c_char_p]),
        "bsg_measure_copy_bandwidth": (C.c_int, [C.c_size_t, C.c_int, pd, pd]),
        "bsg_measure_bandwidth": (C.c_int, [C.c_int, C.c_size_t, C.c_int, pd, pd]),
        "bsg_measure_valu_rate": (C.c_int, [C.c_int, C.c_int, pd, pd]),
        "bsg_measure_sampler_rate": (C.c_int, [C.c_int, pd, pd]),
        "bsg_coef_means": (C.c_int, [vp, C.c_int, pd]),
        "bsg_coef_means_noalloc": (C.c_int, [vp, C.c_int, pd]),
        "bsg_loglik": (C.c_int, [vp, C.c_int, C.c_int, pd]),
        "bsg_nifti_resolve": (C.c_int, [C.c_char_p, C.c_char_p, C.c_size_t]),
        "bsg_nifti_info": (C.c_int, [C.c_char_p, pi, pi]),
        "bsg_nifti_read": (C.c_int, [C.c_char_p, pd, C.c_size_t]),
        "bsg_nifti_write_f64": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, pd]),
        "bsg_design_read": (C.c_int, [C.c_char_p, C.c_int, pi, pf, C.c_char_p, C.c_char_p]),
        "bsg_autotune": (C.c_int, [vp, C.c_int, pi, pd]),
        "bsg_set_lane_split": (C.c_int, [vp, C.c_int]),
        "bsg_get_lane_split": (C.c_int, [vp, pi]),
        "bsg_set_overlap": (C.c_int, [vp, C.c_int]),
        "bsg_test_overlap_wait": (C.c_int, [vp, C.c_int, C.c_int]),
        "bsg_get_overlap": (C.c_int, [vp, pi]),
        "bsg_set_persistent": (C.c_int, [vp, C.c_int]),
        "bsg_get_persistent": (C.c_int, [vp, pi]),
        "bsg_set_persistent_lags": (C.c_int, [vp, C.c_double, C.c_double, C.c_double]),
        "bsg_batch_create": (C.c_int, [C.c_int, C.POINTER(vp), C.POINTER(vp)]),
        "bsg_batch_destroy": (C.c_int, [vp]),
        "bsg_batch_sweep": (C.c_int, [vp, C.c_int]),
        "bsg_batch_sweep_accumulate": (C.c_int, [vp, C.c_int, C.c_int]),
        "bsg_batch_sweep_graph": (C.c_int, [vp, C.c_int, C.c_int]),
        "bsg_batch_set_launch_mode": (C.c_int, [vp, C.c_int]),
        "bsg_batch_get_launch_mode": (C.c_int, [vp, pi]),
        "bsg_kiss": (C.c_double, [C.POINTER(C.c_ulong)]),
        "bsg_devseed": (C.c_uint64, [C.POINTER(C.c_ulong)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise BsgError(lib().bsg_last_error().decode(errors="replace") + f" (status {rc})")


class Layout:
    """The reference's mask layout (hostIdx, hostIdxSC, red/black INDX lists)."""

    def __init__(self, mask, nrow, ncol, ndepth):
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8).reshape(-1))
        if m.size != (nrow + 2) * (ncol + 2) * (ndepth + 2):
            raise ValueError("mask must be the padded (nrow+2)(ncol+2)(ndepth+2) grid")
        self._h = C.c_void_p()
        self.nrow, self.ncol, self.ndepth = nrow, ncol, ndepth
        _check(lib().bsg_layout_create(_ptr(m, C.c_uint8), nrow, ncol, ndepth, C.byref(self._h)))
        cnt = (C.c_int * 4)()
        _check(lib().bsg_layout_counts(self._h, cnt))
        self.totvox, self.totvoxp, self.n0, self.n1 = list(cnt)

    def export(self):
        nsize = (self.nrow + 2) * (self.ncol + 2) * (self.ndepth + 2)
        out = dict(host_idx=np.empty(nsize, np.int32), host_idx_sc=np.empty(nsize, np.int32),
                   vox0=np.empty(self.n0, np.int32), vox1=np.empty(self.n1, np.int32),
                   nbrs0=np.empty(self.n0, np.uint8), nbrs1=np.empty(self.n1, np.uint8),
                   nbr6=np.empty(self.totvox * 6, np.int32))
        _check(lib().bsg_layout_export(
            self._h, _ptr(out["host_idx"], C.c_int), _ptr(out["host_idx_sc"], C.c_int),
            _ptr(out["vox0"], C.c_int), _ptr(out["vox1"], C.c_int), _ptr(out["nbrs0"], C.c_uint8),
            _ptr(out["nbrs1"], C.c_uint8), _ptr(out["nbr6"], C.c_int)))
        return out

    def close(self):
        if self._h:
            lib().bsg_layout_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_opts():
    o = _Opts()
    lib().bsg_opts_default(C.byref(o))
    return o


class Sampler:
    """One sampler context: device state for nchains chains over one data set.

    The options are bsg_opts' (include/bsglmm.h). X'Z forms: xtz_mode 0, an
    FP32 FMA kernel reading Z; 1, BF16 hi + lo on the matrix cores reading Z
    (the precision form with Z stored); 2, FP32 FMA inside the latent draws;
    3, BF16 hi + lo on the matrix cores inside the latent draws. Modes 2 and 3
    keep no Z with z_store=2 (mode 3 requires it)."""

    KERNELS = ("latent", "xtz", "coef0", "coef1", "pairs", "wishart")

    def __init__(self, layout, covar, data, wm=None, seed=0, nchains=1, xtz_mode=0,
                 df_offset=19, trunc_eps=1e-6, pivot_min=1e-6, z_store=0, latent_method=0):
        covar = np.ascontiguousarray(covar, dtype=np.float32)
        data = np.ascontiguousarray(data, dtype=np.uint8)
        self.P, self.N = covar.shape
        self.layout = layout
        self.V, self.VP = layout.totvox, layout.totvoxp
        if data.shape != (self.N, self.V):
            raise ValueError(f"data must be uint8[N={self.N}][TOTVOX={self.V}]")
        o = default_opts()
        o.seed, o.nchains, o.xtz_mode = int(seed), int(nchains), int(xtz_mode)
        o.df_offset, o.trunc_eps, o.pivot_min = int(df_offset), float(trunc_eps), float(pivot_min)
        o.z_store = int(z_store)
        o.latent_method = int(latent_method)
        self.nchains = int(nchains)
        wmp = None
        if wm is not None:
            self._wm = np.ascontiguousarray(wm, dtype=np.float32)
            wmp = _ptr(self._wm, C.c_float)
        self._h = C.c_void_p()
        _check(lib().bsg_create(layout._h, self.N, self.P, _ptr(covar, C.c_float), _ptr(data, C.c_uint8),
                                wmp, C.byref(o), C.byref(self._h)))

    # ---- sweep steps ----
    def sweep(self, n=1):
        _check(lib().bsg_sweep(self._h, int(n)))

    def sweep_graph(self, n, per_graph=8):
        """n sweeps replayed from a HIP graph of per_graph sweeps (same results as sweep)."""
        _check(lib().bsg_sweep_graph(self._h, int(n), int(per_graph)))

    def prepare_graph(self, per_graph):
        """Capture and instantiate the graph of per_graph sweeps that
        sweep_graph uses, without running a sweep."""
        _check(lib().bsg_sweep_graph_prepare(self._h, int(per_graph)))

    def sweep_profiled(self, n):
        ms = np.zeros(6, dtype=np.float64)
        _check(lib().bsg_sweep_profiled(self._h, int(n), _ptr(ms, C.c_double)))
        return dict(zip(self.KERNELS, ms))

    def sweep_timed(self, n):
        """n sweeps; returns (summed latent-kernel ms, total ms) from HIP events
        (overlapped sweeps: the summed spans of each sweep's two launches)."""
        lat, tot = C.c_double(), C.c_double()
        _check(lib().bsg_sweep_timed(self._h, int(n), C.byref(lat), C.byref(tot)))
        return lat.value, tot.value

    # ---- posterior accumulation ----
    def accum_reset(self, ntypes=1, nbatches=50, batch_len=1):
        self._ntypes = int(ntypes)
        _check(lib().bsg_accum_reset(self._h, int(ntypes), int(nbatches), int(batch_len)))

    def accum_add(self):
        _check(lib().bsg_accum_add(self._h))

    def sweep_accumulate(self, n, save_every=1):
        _check(lib().bsg_sweep_accumulate(self._h, int(n), int(save_every)))

    def accum_get(self, chain=0):
        """dict(ndraws, mean[P][V], var[P][V], prb[T][V], bwm[V], mcse[P+T][V]) in mask-rank order."""
        t = C.c_int()
        _check(lib().bsg_accum_shape(self._h, C.byref(t), None, None))
        T = t.value
        out = dict(mean=np.empty((self.P, self.V)), var=np.empty((self.P, self.V)),
                   prb=np.empty((max(T, 1), self.V)), bwm=np.empty(self.V),
                   mcse=np.empty((self.P + T, self.V)))
        n = C.c_int()
        _check(lib().bsg_accum_get(self._h, int(chain), C.byref(n), _ptr(out["mean"], C.c_double),
                                   _ptr(out["var"], C.c_double), _ptr(out["prb"], C.c_double),
                                   _ptr(out["bwm"], C.c_double), _ptr(out["mcse"], C.c_double)))
        out["ndraws"] = n.value
        out["prb"] = out["prb"][:T]
        return out

    def update_latent(self):
        _check(lib().bsg_update_latent(self._h))

    def compute_xtz(self):
        _check(lib().bsg_compute_xtz(self._h))

    def update_wm_coef(self):
        _check(lib().bsg_update_wm_coef(self._h))

    def update_coef(self, colour):
        _check(lib().bsg_update_coef(self._h, int(colour)))

    def update_car_prec(self):
        _check(lib().bsg_update_car_prec(self._h))

    def synchronize(self):
        _check(lib().bsg_synchronize(self._h))

    # ---- state, reference layouts ----
    def get_spat_coef(self, chain=0):
        out = np.empty((self.P, self.VP), np.float32)
        _check(lib().bsg_get_spat_coef(self._h, chain, _ptr(out, C.c_float)))
        return out

    def set_spat_coef(self, sc, chain=0):
        sc = np.ascontiguousarray(sc, dtype=np.float32)
        assert sc.shape == (self.P, self.VP)
        _check(lib().bsg_set_spat_coef(self._h, chain, _ptr(sc, C.c_float)))

    def get_z(self, chain=0):
        out = np.empty((self.N, self.V), np.float32)
        _check(lib().bsg_get_z(self._h, chain, _ptr(out, C.c_float)))
        return out

    def get_xtz(self, chain=0):
        out = np.empty((self.P, self.V), np.float32)
        _check(lib().bsg_get_xtz(self._h, chain, _ptr(out, C.c_float)))
        return out

    def set_xtz(self, m, chain=0):
        m = np.ascontiguousarray(m, dtype=np.float32)
        _check(lib().bsg_set_xtz(self._h, chain, _ptr(m, C.c_float)))

    def autotune(self, reps=10):
        """Lane split G of the Z-free sweep chosen by timing whole sweeps (the
        chain state is restored); (G, ms per sweep for G = 1, 2, 4, 8)."""
        g = C.c_int()
        ms = np.empty(4, np.float64)
        _check(lib().bsg_autotune(self._h, int(reps), C.byref(g), _ptr(ms, C.c_double)))
        return g.value, ms

    @property
    def lane_split(self):
        """Lanes per voxel G of the Z-free latent kernel the next sweep uses."""
        g = C.c_int()
        _check(lib().bsg_get_lane_split(self._h, C.byref(g)))
        return g.value

    @lane_split.setter
    def lane_split(self, g):
        _check(lib().bsg_set_lane_split(self._h, int(g)))

    @property
    def overlap(self):
        """True when the next sweep runs overlapped: two launches per sweep,
        each colour's coefficient update beside the other colour's latent
        draws (four lanes per voxel, no WM image). The chains are
        the same either way. Assign False to run the serial order."""
        on = C.c_int()
        _check(lib().bsg_get_overlap(self._h, C.byref(on)))
        return bool(on.value)

    @overlap.setter
    def overlap(self, on):
        _check(lib().bsg_set_overlap(self._h, 1 if on else 0))

    @property
    def persistent(self):
        """True when the next sweep() runs as one persistent launch (an
        overlapped one-chain context with P <= 6, asked for by assigning
        True): resident blocks take the sweeps' latent tiles, colour updates
        and CAR-precision draws from a queue ordered by what each reads. The
        chains are the same either way."""
        on = C.c_int()
        _check(lib().bsg_get_persistent(self._h, C.byref(on)))
        return bool(on.value)

    @persistent.setter
    def persistent(self, on):
        _check(lib().bsg_set_persistent(self._h, 1 if on else 0))

    def set_persistent_lags(self, latent, update, draw):
        """Queue lags of the persistent sweep, as fractions of its resident
        block count: how far behind its inputs each update and draw is
        placed. They change the speed only."""
        _check(lib().bsg_set_persistent_lags(self._h, float(latent), float(update), float(draw)))

    def coef_means(self, chain=0):
        """Plane means of the coefficients over the mask (alphaMean)."""
        out = np.empty(self.P, np.float64)
        _check(lib().bsg_coef_means(self._h, chain, _ptr(out, C.c_double)))
        return out

    def loglik(self, ntypes=1, chain=0):
        """Per-subject log-likelihood under each subtype, float64[N][ntypes]."""
        out = np.empty((self.N, ntypes), np.float64)
        _check(lib().bsg_loglik(self._h, chain, int(ntypes), _ptr(out, C.c_double)))
        return out

    def get_wm_coef(self, chain=0):
        v = C.c_float()
        _check(lib().bsg_get_wm_coef(self._h, chain, C.byref(v)))
        return v.value

    def set_wm_coef(self, b, chain=0):
        _check(lib().bsg_set_wm_coef(self._h, chain, C.c_float(b)))

    def get_prec(self, chain=0):
        out = np.empty((self.P, self.P), np.float32)
        _check(lib().bsg_get_prec(self._h, chain, _ptr(out, C.c_float)))
        return out

    def set_prec(self, prec, chain=0):
        prec = np.ascontiguousarray(prec, dtype=np.float32)
        _check(lib().bsg_set_prec(self._h, chain, _ptr(prec, C.c_float)))

    def get_pair_sums(self, chain=0):
        out = np.empty((self.P, self.P), np.float64)
        _check(lib().bsg_get_pair_sums(self._h, chain, _ptr(out, C.c_double)))
        return out

    def sweep_index(self, chain=0):
        v = C.c_uint32()
        _check(lib().bsg_get_sweep_index(self._h, chain, C.byref(v)))
        return v.value

    def set_sweep_index(self, sweep):
        _check(lib().bsg_set_sweep_index(self._h, C.c_uint32(sweep)))

    def fallback_count(self):
        v = C.c_uint32()
        _check(lib().bsg_get_fallback_count(self._h, C.byref(v)))
        return v.value

    def checkpoint(self, path):
        """Writes the whole chain state (save_iter, mcmc.cpp:1265-1296, exact)."""
        _check(lib().bsg_checkpoint_write(self._h, os.fsencode(path)))

    def restore(self, path):
        """Reads a checkpoint of the same problem (restart_iter, mcmc.cpp:1298-1336);
        the chains then continue bit for bit."""
        _check(lib().bsg_checkpoint_read(self._h, os.fsencode(path)))

    def close(self):
        if self._h:
            lib().bsg_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_bandwidth(nbytes=1 << 30, reps=20):
    """(GB/s, ms per copy) of a float4 stream copy over nbytes (read + write counted)."""
    g, ms = C.c_double(), C.c_double()
    _check(lib().bsg_measure_copy_bandwidth(C.c_size_t(nbytes), int(reps), C.byref(g), C.byref(ms)))
    return g.value, ms.value


def valu_rate(kind="fma", reps=100):
    """(wave64 VALU instructions per second over the chip, ms per launch) of a
    VALU-bound probe: kind 'fma' (v_fma_f32) or 'trans' (v_exp_f32)."""
    k = {"fma": 0, "trans": 1}[kind]
    w, ms = C.c_double(), C.c_double()
    _check(lib().bsg_measure_valu_rate(k, int(reps), C.byref(w), C.byref(ms)))
    return w.value, ms.value


def sampler_rate(reps=100):
    """(draws per second over the chip, ms per launch) of the latent sampler
    alone: Philox4x32-10 and the inversion draw, no memory traffic."""
    r, ms = C.c_double(), C.c_double()
    _check(lib().bsg_measure_sampler_rate(int(reps), C.byref(r), C.byref(ms)))
    return r.value, ms.value


def stream_bandwidth(kind, nbytes=1 << 30, reps=20):
    """(GB/s, ms per launch) of a stream kernel: kind 'copy', 'write' or 'read'."""
    k = {"copy": 0, "write": 1, "read": 2}[kind]
    g, ms = C.c_double(), C.c_double()
    _check(lib().bsg_measure_bandwidth(k, C.c_size_t(nbytes), int(reps), C.byref(g), C.byref(ms)))
    return g.value, ms.value


def kiss(seed):
    """One KISS draw (randgen.cpp:9-23); seed is a length-3 uint64 array, advanced in place."""
    s = (C.c_ulong * 3)(*[int(x) for x in seed])
    u = lib().bsg_kiss(s)
    seed[:] = list(s)
    return u


def devseed(seed):
    """Philox key from a seed.dat triple, as main.cu:413 derives devseed."""
    s = (C.c_ulong * 3)(*[int(x) for x in seed])
    v = lib().bsg_devseed(s)
    seed[:] = list(s)
    return v


class Batch:
    """Independent analyses (Samplers in the Z-free form) swept together:
    each phase of a sweep is one kernel launch per covariate count over every
    analysis and chain with that count (the latent phase one more for the
    members at four lanes per voxel)."""

    def __init__(self, samplers):
        self.samplers = list(samplers)
        arr = (C.c_void_p * len(self.samplers))(*[S._h.value for S in self.samplers])
        self._h = C.c_void_p()
        _check(lib().bsg_batch_create(len(self.samplers), arr, C.byref(self._h)))

    def sweep(self, n=1):
        _check(lib().bsg_batch_sweep(self._h, int(n)))

    def sweep_accumulate(self, n, save_every=1):
        _check(lib().bsg_batch_sweep_accumulate(self._h, int(n), int(save_every)))

    def sweep_graph(self, n, per_graph=8):
        """n batch sweeps replayed from a HIP graph of per_graph sweeps (same results as sweep)."""
        _check(lib().bsg_batch_sweep_graph(self._h, int(n), int(per_graph)))

    @property
    def one_launch_per_phase(self):
        """True when a batch sweep is four launches over every covariate count."""
        v = C.c_int()
        _check(lib().bsg_batch_get_launch_mode(self._h, C.byref(v)))
        return bool(v.value)

    def set_launch_mode(self, mode):
        """0: chosen from the draws per sweep; 1: one launch per phase; 2: one per phase and P."""
        _check(lib().bsg_batch_set_launch_mode(self._h, int(mode)))

    def close(self):
        if self._h:
            lib().bsg_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- file formats of the reference program (include/bsglmm_io.h) ----
def nifti_resolve(path):
    buf = C.create_string_buffer(4096)
    _check(lib().bsg_nifti_resolve(os.fsencode(path), buf, 4096))
    return buf.value.decode()


def nifti_info(path):
    """(dims (nx, ny, nz), NIfTI datatype code) of volume 1."""
    dims = (C.c_int * 3)()
    dt = C.c_int()
    _check(lib().bsg_nifti_info(os.fsencode(path), dims, C.byref(dt)))
    return tuple(dims), dt.value


def nifti_read(path):
    """Volume 1 as float64[nz][ny][nx] (x fastest in memory)."""
    (nx, ny, nz), _ = nifti_info(path)
    out = np.empty((nz, ny, nx), np.float64)
    _check(lib().bsg_nifti_read(os.fsencode(path), _ptr(out, C.c_double), out.size))
    return out


def nifti_write(path, vol):
    """Writes float64[nz][ny][nx] as a FLOAT64 NIfTI-1 (.nii, or .nii.gz compressed)."""
    v = np.ascontiguousarray(vol, dtype=np.float64)
    nz, ny, nx = v.shape
    _check(lib().bsg_nifti_write_f64(os.fsencode(path), nx, ny, nz, _ptr(v, C.c_double)))


def design_read(path, ncovar):
    """(covar float32[ncovar][N], covariate names, image paths) of a design file."""
    n = C.c_int()
    _check(lib().bsg_design_read(os.fsencode(path), int(ncovar), C.byref(n), None, None, None))
    covar = np.empty((ncovar, n.value), np.float32)
    names = C.create_string_buffer(64 * ncovar)
    imgs = C.create_string_buffer(512 * n.value)
    _check(lib().bsg_design_read(os.fsencode(path), int(ncovar), C.byref(n), _ptr(covar, C.c_float), names, imgs))
    nm = [names.raw[64 * c:64 * (c + 1)].split(b"\0")[0].decode() for c in range(ncovar)]
    im = [imgs.raw[512 * i:512 * (i + 1)].split(b"\0")[0].decode() for i in range(n.value)]
    return covar, nm, im
