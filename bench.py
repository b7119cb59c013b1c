"""Benchmark: Gibbs sweeps/s of the BSGLMM probit sweep on one MI355X.

Workload (the configuration the metric is quoted on): grid 91x109x91 with an
ellipsoidal-shell mask of V = 230,906 voxels, N = 250 subjects, P = 4
covariates, one chain, FP32 state, FP32-FMA X'Z, synthetic data from the
model (bsglmm_amd.synth). A step is one full sweep: latent draws, X'Z,
checkerboard coefficient updates (two colours), FP64 pair sums and the device
Wishart draw. All inputs are resident in HBM before timing starts.

The headline configuration forms X'Z inside the latent kernel (FP32 FMA);
the same sweep with a separate X'Z kernel re-reading Z is timed beside it.
Prints one JSON line with the dominant kernel's roofline (latent draws with
fused X'Z: algorithmic bytes N*V*(4 + 1/8) + 8*P*V per launch against
8 TB/s), a measured copy-kernel bandwidth, and the reference CPU path timed on
this host for a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gibbs sweeps/s + latent draws/s (N=250,V≈230k,P=4); ESS/s batched vs plain"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--cpu-sweeps", type=int, default=2, help="timed reference CPU sweeps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "latent_c2_traffic.json"))
    return ap.parse_args()


def cpu_baseline(d, nsweeps):
    """Reference CPU path (oracle/_ref: bsglmm/covar.cpp, randgen.cpp, cholesky.cpp
    compiled unchanged) timed for a bounded number of sweeps on one pinned core."""
    import ctypes as C

    import numpy as np

    import oracle
    if not oracle.ref_available():
        return None
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    covar = d["covar"]
    P, N = covar.shape
    V, VP = lay["totvox"], lay["totvoxp"]
    oracle.RefCPU(lay, covar, oracle.xtx_reference(covar))
    R = oracle.ref()
    z = np.zeros((N, V), np.float32)
    data = np.ascontiguousarray(d["data"])
    wm = np.ones(V, np.float32)
    sc = np.zeros((P, VP), np.float32)
    prec = np.eye(P, dtype=np.float32)
    seed = oracle.seed_array([5118772087888418775, 289433066200759127, 1964621368])
    f, u8 = C.c_float, C.c_uint8
    args = (z.ctypes.data_as(C.POINTER(f)), data.ctypes.data_as(C.POINTER(u8)), wm.ctypes.data_as(C.POINTER(f)),
            sc.ctypes.data_as(C.POINTER(f)), covar.ctypes.data_as(C.POINTER(f)),
            prec.ctypes.data_as(C.POINTER(f)), seed)
    old = os.sched_getaffinity(0)
    try:
        os.sched_setaffinity(0, {min(old)})
        R.ref_sweep(*args)  # warm-up sweep
        t0 = time.perf_counter()
        for _ in range(nsweeps):
            R.ref_sweep(*args)
        dt = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, old)
    return {"value": nsweeps / dt, "unit": "sweeps/s", "cores": 1, "kind": "reference",
            "sample": f"{nsweeps} full sweeps of the same workload after 1 warm-up sweep, "
                      f"reference updateZ/updateSpatCoef/updateSpatCoefPrec on 1 pinned core "
                      f"of {os.cpu_count()}; {dt:.1f} s",
            "latent_draws_per_s": N * V * nsweeps / dt}


def time_mode(L, d, mode, steps, warmup):
    """Timed region: `steps` sweeps (CAR-precision draw overlapped with the next
    latent draw), HIP events around every latent kernel on its stream. Then a
    short serial pass with events around every kernel for the breakdown."""
    from bsglmm_amd import Sampler
    S = Sampler(L, d["covar"], d["data"], seed=0x9E3779B97F4A7C15, xtz_mode=mode)
    S.sweep(warmup)
    S.synchronize()
    t0 = time.perf_counter()
    lat_ms, span_ms = S.sweep_timed(steps)
    S.synchronize()
    wall = time.perf_counter() - t0
    kms = S.sweep_profiled(min(steps, 50))
    n_prof = min(steps, 50)
    S.close()
    return wall, lat_ms, span_ms, {k: v / n_prof for k, v in kms.items()}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {a.gpus} is not supported "
              "(the sweep shards over voxel slabs by design but this build measures one GPU)",
              file=sys.stderr)
        sys.exit(2)
    from bsglmm_amd import Layout, copy_bandwidth, synth

    d = synth.generate(a.config)
    nx, ny, nz = d["grid"]
    L = Layout(d["mask"], nx, ny, nz)
    P, N = d["covar"].shape
    V = L.totvox
    copy_gbs, copy_ms = copy_bandwidth(1 << 30, 20)

    # plain single chain, X'Z fused into the latent kernel (FP32 FMA)
    wall, lat_ms, span_ms, kms = time_mode(L, d, 2, a.steps, a.warmup)
    # plain single chain, separate FP32 X'Z kernel reading Z (the layout of the reference sweep)
    wall_sep, lat_ms_sep, span_sep, kms_sep = time_mode(L, d, 0, a.steps, a.warmup)

    value = a.steps / wall
    gpu_ms = span_ms / a.steps
    lat_bytes = N * V * (4 + 1 / 8) + 8 * P * V
    lat_s = lat_ms / a.steps / 1e3
    achieved = lat_bytes / lat_s / 1e9
    traffic = None
    if os.path.exists(a.traffic):
        with open(a.traffic) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    sweep_bytes_fused = N * V * (4 + 1 / 8) + V * (24 * P + 5)
    sweep_bytes_sep = N * V * (4 + 4 + 1 / 8) + V * (24 * P + 5)
    gpu_ms_sep = span_sep / a.steps
    out = {
        "metric": METRIC, "value": value, "unit": "sweeps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{a.config}: grid {nx}x{ny}x{nz} ellipsoidal-shell mask, plain single chain, "
                               f"FP32 state, FP32-FMA X'Z fused into the latent kernel",
                   "V": V, "N": N, "P": P, "chains": 1},
        "latent_draws_per_s": value * N * V,
        "latent_kernel_ms": lat_ms / a.steps,
        "kernel_ms_serial_breakdown": kms,
        "sweep_gpu_ms": gpu_ms,
        "sweep_hbm_frac": sweep_bytes_fused / (gpu_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
        "copy_kernel_gbs": copy_gbs,
        "roofline": {"kernel": "k_latent (X'Z fused)", "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes_per_launch": lat_bytes,
                     "frac_of_copy": achieved / copy_gbs},
        "plain_separate_xtz": {
            "value": a.steps / wall_sep, "unit": "sweeps/s", "ms_per_step": 1e3 * wall_sep / a.steps,
            "kernel_ms_serial_breakdown": kms_sep,
            "sweep_gpu_ms": gpu_ms_sep,
            "sweep_hbm_frac": sweep_bytes_sep / (gpu_ms_sep / 1e3) / 1e9 / HBM_PEAK_GBS,
            "latent_gbs": (N * V * (4 + 1 / 8) + 4 * P * V) / (lat_ms_sep / a.steps / 1e3) / 1e9,
            "xtz_gbs": (4 * N * V + 4 * P * V) / (kms_sep["xtz"] / 1e3) / 1e9},
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(d, a.cpu_sweeps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
