"""Benchmark: Gibbs sweeps/s of the BSGLMM probit sweep on one MI355X.

Workload (the configuration the metric is quoted on): grid 91x109x91 with an
ellipsoidal-shell mask of V = 230,906 voxels, N = 250 subjects, P = 4
covariates, one chain, FP32 state, FP32-FMA X'Z, synthetic data from the
model (bsglmm_amd.synth). A step is one full sweep: latent draws, X'Z,
checkerboard coefficient updates (two colours), FP64 pair sums and the device
Wishart draw. All inputs are resident in HBM before timing starts.

Headline form: X'Z is formed inside the latent kernel from draws held in
registers (FP32 FMA), Z is not written, and the sweep is overlapped: two
launches of k_overlap per sweep, each drawing one colour's latent variates
(per-draw means x_i'b_j on FP32 MFMAs) beside the other colour's coefficient
update, with the CAR-precision draw inside the first. Beside it, in the same
run: the same Z-free sweep in the serial order (k_latent_vl_mfm, then the two
colour updates; the same chain bit for bit), the fused kernel that also
stores Z, and the sweep with a separate X'Z kernel re-reading Z (FP32 FMA,
and BF16 hi+lo on MFMA); the C1 configuration's full 2,000-iteration run
(overlapped, and replayed from HIP graphs of 50 sweeps) and the C3 precision
form (N = 1,000, P = 8: BF16 hi+lo MFMA X'Z on stored Z, beside the Z-free
FP32 form); the 32-chain batch (C4: ESS/s and R-hat against the plain chain)
and the twelve uneven analyses (C5).

Prints one JSON line with the dominant kernel's roofline (its algorithmic
bytes against 8 TB/s, its counted HBM bytes, and the limiter it actually
meets: VALU issue, from the VALU count of its rocprofv3 profile over the
launch time measured here, against the VALU issue rate a probe kernel
measures in the same run; and against the draw rate of the latent sampler
alone, a probe without memory traffic), the effective sweep-level HBM fraction (the
unfused sweep's bytes N V (8 + 1/8) + V (24 P + 5), which credits Z traffic
this form never moves), a measured copy-kernel bandwidth, and the reference
CPU path: C2 for 3 warm-up + 20 timed sweeps and C1 for its full 2,000
iterations, each on one pinned host core of its own while the GPU work runs.
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
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--cpu-sweeps", type=int, default=20, help="timed reference CPU sweeps at C2")
    ap.add_argument("--cpu-warmup", type=int, default=3, help="untimed reference CPU sweeps at C2 first")
    ap.add_argument("--cpu-c1-sweeps", type=int, default=2000, help="reference CPU sweeps at C1 (the full run)")
    ap.add_argument("--cpu-leg", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true",
                    help="time the headline sweeps as replays of a HIP graph prepared before the warm-up "
                         "(at 20 sweeps no faster than direct launches on this machine)")
    ap.add_argument("--persistent", action="store_true",
                    help="time the headline sweeps in the persistent form (one k_persist launch per call; "
                         "the same chain, slower on this machine than the overlapped default)")
    ap.add_argument("--config-lines", type=int, default=1, help="also time the C1 and C3 configurations")
    ap.add_argument("--c1-sweeps", type=int, default=2000, help="timed sweeps of the C1 line (its full run)")
    ap.add_argument("--c3-sweeps", type=int, default=200, help="timed sweeps of each C3 form")
    ap.add_argument("--batched-chains", type=int, default=32,
                    help="chains of the batched form (0 skips the ESS comparison)")
    ap.add_argument("--batched-iters", type=int, default=5000, help="iterations per chain, burn-in included")
    ap.add_argument("--batched-burnin", type=int, default=1000)
    ap.add_argument("--uneven-steps", type=int, default=50,
                    help="batch sweeps of the 12-analysis uneven batch (0 skips it)")
    ap.add_argument("--profile", default=os.path.join(ROOT, "profiles", "k_overlap_c2.json"),
                    help="measured traffic and VALU counts of the dominant kernel (rocprofv3 summary)")
    return ap.parse_args()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(name, warm, timed, core):
    """The reference CPU path (oracle/_ref: bsglmm/covar.cpp, randgen.cpp and
    cholesky.cpp compiled unchanged) on one pinned core, in the order of
    mcmc.cpp:400-428: updateZ, updateBeta, updateSpatCoef, updateSpatCoefPrec.
    updateBeta lives in mcmc.cpp, which cannot be compiled without CUDA; its
    statistic runs as the oracle's restatement (orc_cpu_wm_stat, float Kahan
    sums in raster order as mcmc.cpp:1099-1136) and its normal draw through
    the reference's own generator. Returns sweeps/s over `timed` sweeps after
    `warm` untimed ones."""
    import ctypes as C

    import numpy as np

    import oracle
    from bsglmm_amd import synth
    os.sched_setaffinity(0, {core})
    if not oracle.ref_available():
        return None
    d = synth.generate(name)
    nx, ny, nz = d["grid"]
    lay = oracle.layout(d["mask"], nx, ny, nz)
    covar = np.ascontiguousarray(d["covar"], np.float32)
    P, N = covar.shape
    V, VP = lay["totvox"], lay["totvoxp"]
    oracle.RefCPU(lay, covar, oracle.xtx_reference(covar))
    R = oracle.ref()
    O = oracle.Oracle(lay, covar, d["data"])
    z = np.zeros((N, V), np.float32)
    data = np.ascontiguousarray(d["data"])
    wm = np.ones(V, np.float32)
    sc = np.zeros((P, VP), np.float32)
    prec = np.eye(P, dtype=np.float32)
    seed = oracle.seed_array([5118772087888418775, 289433066200759127, 1964621368])
    f, u8 = C.c_float, C.c_uint8
    pz, pd, pw = z.ctypes.data_as(C.POINTER(f)), data.ctypes.data_as(C.POINTER(u8)), wm.ctypes.data_as(C.POINTER(f))
    ps, pc, pp = sc.ctypes.data_as(C.POINTER(f)), covar.ctypes.data_as(C.POINTER(f)), prec.ctypes.data_as(C.POINTER(f))
    b = 0.0
    t_beta = 0.0

    def sweep():
        nonlocal b, t_beta
        R.ref_update_z(pz, pd, b, pw, ps, pc, seed)
        t = time.perf_counter()
        m, v = O.cpu_wm_stat(sc, z, 0.0, 1e-6)  # prior_mean_beta 0, prior_prec_beta 1e-6 (mcmc.cpp:81)
        b = float(np.float32(m + np.sqrt(v) * R.ref_snorm(seed)))
        t_beta += time.perf_counter() - t
        R.ref_update_spat_coef(pc, ps, pp, pz, b, pw, seed)
        R.ref_update_spat_coef_prec(pp, ps, seed)

    for _ in range(warm):
        sweep()
    t_beta = 0.0
    t0 = time.perf_counter()
    for _ in range(timed):
        sweep()
    dt = time.perf_counter() - t0
    return {"config": name, "V": V, "N": N, "P": P, "warmup_sweeps": warm, "timed_sweeps": timed,
            "seconds": dt, "sweeps_per_s": timed / dt, "latent_draws_per_s": N * V * timed / dt,
            "update_beta_share": t_beta / dt, "core": core}


def start_cpu_legs(legs):
    """Each (config, warm-up sweeps, timed sweeps) leg in a child process pinned
    to a core of its own (the last cores of this process's set, which this
    process then leaves to them). Started before anything here touches the GPU."""
    import subprocess
    cores = sorted(os.sched_getaffinity(0))
    procs = []
    for k, (name, warm, timed) in enumerate(legs):
        core = cores[-1 - k] if len(cores) > k + 1 else cores[-1]
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-leg", f"{name}:{warm}:{timed}:{core}"]
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
    mine = set(cores[:-len(legs)]) if len(cores) > len(legs) else set(cores)
    os.sched_setaffinity(0, mine)
    return procs


def finish_cpu_legs(procs, timeout=900):
    out = {}
    for p in procs:
        try:
            so, se = p.communicate(timeout=timeout)
        except Exception:
            p.kill()
            so, se = p.communicate()
        for line in so.splitlines():
            if line.startswith("{"):
                r = json.loads(line)
                if r:
                    out[r["config"]] = r
    if "c2" not in out:
        return None
    c2 = out["c2"]
    res = {"value": c2["sweeps_per_s"], "unit": "sweeps/s", "cores": 1, "kind": "reference",
           "sample": (f"{c2['warmup_sweeps']} warm-up + {c2['timed_sweeps']} timed full sweeps of the same workload "
                      f"(c2) in {c2['seconds']:.1f} s on 1 pinned core of {os.cpu_count()} ({cpu_model()}); "
                      f"reference updateZ/updateSpatCoef/updateSpatCoefPrec compiled unchanged, updateBeta's "
                      f"statistic restated (mcmc.cpp needs CUDA), "
                      f"{100 * c2['update_beta_share']:.1f} % of the sweep"),
           "latent_draws_per_s": c2["latent_draws_per_s"], "cpu_model": cpu_model(), "host_cores": os.cpu_count()}
    if "c1" in out:
        res["c1"] = out["c1"]
    return res


def time_mode(L, d, mode, zstore, steps, warmup, overlap=True, settle=None, profile_sweeps=50, graph=False,
              persistent=False):
    """Timed region: `steps` sweeps after `warmup` untimed direct ones,
    bracketed by stream synchronisations: direct launches (bsg_sweep), or
    with graph=True replays of a HIP graph of min(steps, 256) sweeps
    (bsg_sweep_graph; captured and uploaded before any GPU work here, so the
    timed call only replays it), and then `steps` direct sweeps timed the
    same way for comparison. `settle` (optional) runs just before the
    warm-up, after the context is built, so that the GPU is busy and at its
    working clock when the few warm-up sweeps start (bench.py's
    copy-bandwidth, VALU-rate and sampler-rate probes). Then, in the same run, `steps` more sweeps
    with HIP events around every latent launch on the stream it runs on
    (kernel duration and sweep span; the event packets cost a few
    microseconds per sweep, so they stay out of the timed region), and a
    short serial pass with events around every kernel for the breakdown.
    Overlapped sweeps (where the form allows them and overlap is True): the
    events bracket each sweep's two launches."""
    from bsglmm_amd import Sampler
    S = Sampler(L, d["covar"], d["data"], seed=0x9E3779B97F4A7C15, xtz_mode=mode, z_store=zstore)
    if overlap is not None:
        S.overlap = overlap
    ran_overlapped = S.overlap
    if persistent:
        S.persistent = True  # bsg_sweep: one k_persist launch per call (same chain)
        if not S.persistent:
            raise RuntimeError("the persistent sweep does not support this workload")
    per_graph = min(steps, 256)
    if graph:
        S.prepare_graph(per_graph)
    S.synchronize()
    extra = settle() if settle else None
    S.sweep(warmup)
    S.synchronize()
    t0 = time.perf_counter()
    if graph:
        S.sweep_graph(steps, per_graph)
    else:
        S.sweep(steps)
    S.synchronize()
    wall = time.perf_counter() - t0
    direct_wall = None
    if graph:
        t0 = time.perf_counter()
        S.sweep(steps)
        S.synchronize()
        direct_wall = time.perf_counter() - t0
    n_ev = max(steps, 200)
    lat_ms, span_ms = S.sweep_timed(n_ev)
    S.synchronize()
    n_prof = min(steps, profile_sweeps)
    kms = S.sweep_profiled(n_prof) if n_prof > 0 else {}
    S.close()
    return {"wall": wall, "lat_ms": lat_ms / n_ev, "span_ms": span_ms / n_ev,
            "kms": {k: v / n_prof for k, v in kms.items()}, "overlapped": ran_overlapped, "settle": extra,
            "graph_group": per_graph if graph else 0, "direct_wall": direct_wall}


def config_line(name, forms, steps, warmup):
    """Sweeps/s of configuration `name` in each (label, xtz_mode, z_store,
    overlap, graph group) form: `steps` sweeps after `warmup`, timed by the
    host clock around the stream (graph group > 0: bsg_sweep_graph; overlap
    None: the context's default)."""
    from bsglmm_amd import Layout, Sampler, synth
    d = synth.generate(name)
    L = Layout(d["mask"], *d["grid"])
    P, N = d["covar"].shape
    V = L.totvox
    out = {"config": name, "V": V, "N": N, "P": P, "chains": 1, "timed_sweeps": steps, "forms": {}}
    for label, mode, zs, ov, grp in forms:
        S = Sampler(L, d["covar"], d["data"], seed=0x5EED0C0F, xtz_mode=mode, z_store=zs)
        if ov is not None:
            S.overlap = ov
        S.sweep(warmup)
        S.synchronize()
        t0 = time.perf_counter()
        if grp:
            S.sweep_graph(steps, grp)
        else:
            S.sweep(steps)
        S.synchronize()
        wall = time.perf_counter() - t0
        lat_ms, _ = S.sweep_timed(min(steps, 200))
        r = {"xtz_mode": mode, "z_store": zs, "overlapped": S.overlap, "graph_group": grp,
             "sweeps_per_s": steps / wall, "us_per_sweep": 1e6 * wall / steps,
             "latent_draws_per_s": N * V * steps / wall,
             "latent_kernel_us": 1e3 * lat_ms / min(steps, 200)}
        if mode < 2:
            r["kernel_us_serial_breakdown"] = {k: 1e3 * v / 20 for k, v in S.sweep_profiled(20).items()}
        out["forms"][label] = r
        S.close()
    return out


def chain_summaries(L, d, chains, iters, burnin, nbatches=40, xtz_mode=2):
    """`chains` chains of the Z-free fused sweep in one context (every kernel
    launch covers all chains); after burn-in, every draw goes to the device
    accumulators. Returns the sampling wall time (burn-in excluded), the
    per-chain batch-means ESS of every coefficient map (var / mcse^2)
    and the per-voxel R-hat across chains (Gelman-Rubin, chain means and
    variances). Throughput counts chain-sweeps."""
    import numpy as np

    from bsglmm_amd import Sampler
    S = Sampler(L, d["covar"], d["data"], seed=0x5EED5EED12345678, nchains=chains, xtz_mode=xtz_mode, z_store=2)
    t0 = time.perf_counter()
    S.sweep(burnin)
    S.synchronize()
    t_burn = time.perf_counter() - t0
    n = iters - burnin
    S.accum_reset(ntypes=1, nbatches=nbatches, batch_len=n // nbatches)
    t0 = time.perf_counter()
    S.sweep_accumulate(n, 1)
    S.synchronize()
    wall = time.perf_counter() - t0
    P = S.P
    means, vars_, ess = [], [], []
    for c in range(chains):
        o = S.accum_get(c)
        means.append(o["mean"])
        vars_.append(o["var"])
        ess.append(o["var"] / np.maximum(o["mcse"][:P] ** 2, 1e-300))  # var / Var(mean)
    S.close()
    means, vars_, ess = np.array(means), np.array(vars_), np.array(ess)  # chains x P x V
    rhat = None
    if chains > 1:
        W = vars_.mean(0)
        B = n * means.var(0, ddof=1)
        rhat = np.sqrt(((n - 1) / n * W + B / n) / W)
    return {"wall": wall, "burnin_wall": t_burn, "ess": ess, "rhat": rhat, "draws": n}


def batched_vs_plain(L, d, chains, iters, burnin):
    import numpy as np
    plain = chain_summaries(L, d, 1, iters, burnin)
    many = chain_summaries(L, d, chains, iters, burnin)
    P = plain["ess"].shape[1]
    out = {"chains": chains, "iterations": iters, "burnin": burnin, "draws_per_chain": many["draws"],
           "estimator": "batch-means ESS per voxel (40 batches), summed over chains; ESS/s over the sampling "
                        "phase wall time (accumulation included, burn-in excluded)",
           "plain_chain_sweeps_per_s": plain["draws"] / plain["wall"],
           "batched_chain_sweeps_per_s": chains * many["draws"] / many["wall"], "maps": []}
    ratios = []
    for c in range(P):
        ess_plain = plain["ess"][0, c]
        ess_many = many["ess"][:, c].sum(0)
        r_plain = np.median(ess_plain) / plain["wall"]
        r_many = np.median(ess_many) / many["wall"]
        ratios.append(r_many / r_plain)
        m = {"coef": c, "plain_ess_per_s_median": r_plain, "batched_ess_per_s_median": r_many,
             "ratio": r_many / r_plain,
             "plain_ess_median": float(np.median(ess_plain)), "plain_ess_p05": float(np.percentile(ess_plain, 5)),
             "batched_ess_per_chain_median": float(np.median(many["ess"][:, c]))}
        if many["rhat"] is not None:
            rh = many["rhat"][c]
            # across-chain check of the batch-means ESS: Var(chain mean) = B / n
            m["ess_per_chain_from_between_chain_variance_median"] = float(
                np.median(1.0 / np.maximum((rh ** 2 - 1.0 + 1.0 / many["draws"]), 1e-12)))
            m.update({"rhat_median": float(np.median(rh)), "rhat_p99": float(np.percentile(rh, 99)),
                      "rhat_max": float(rh.max()), "rhat_frac_above_1_01": float(np.mean(rh > 1.01))})
        out["maps"].append(m)
    out["ess_per_s_ratio_min_over_maps"] = float(min(ratios))
    # the chains' sweeps as a hipGraph-captured iteration: replays of a prepared
    # graph of 50 sweeps against the same number of direct sweeps (no accumulation)
    from bsglmm_amd import Sampler
    G = Sampler(L, d["covar"], d["data"], seed=0x5EED5EED12345678, nchains=chains, xtz_mode=2, z_store=2)
    G.prepare_graph(50)
    G.sweep(50)
    G.synchronize()
    t0 = time.perf_counter()
    G.sweep_graph(500, 50)
    G.synchronize()
    t_graph = time.perf_counter() - t0
    t0 = time.perf_counter()
    G.sweep(500)
    G.synchronize()
    t_direct = time.perf_counter() - t0
    G.close()
    out["graph_replay"] = {"sweeps": 500, "per_graph": 50, "chain_sweeps_per_s": chains * 500 / t_graph,
                           "direct_chain_sweeps_per_s": chains * 500 / t_direct}
    # the same chains in the mixed-precision form (X'Z as BF16 hi + lo products on
    # the matrix cores inside the draws) against the same plain FP32 chain
    mp = chain_summaries(L, d, chains, iters, burnin, xtz_mode=3)
    mr = [(np.median(mp["ess"][:, c].sum(0)) / mp["wall"]) / (np.median(plain["ess"][0, c]) / plain["wall"])
          for c in range(P)]
    out["mixed_precision"] = {"xtz_mode": 3, "batched_chain_sweeps_per_s": chains * mp["draws"] / mp["wall"],
                              "ess_per_s_ratio_by_map": [float(r) for r in mr],
                              "ess_per_s_ratio_min_over_maps": float(min(mr)),
                              "rhat_max": float(mp["rhat"].max()) if mp["rhat"] is not None else None}
    return out


def uneven_batch(steps, xtz_mode=3, extras=True):
    """Twelve independent analyses (ellipsoid masks of 20,031 to 398,905
    voxels, N from 40 to 1,000, P from 2 to 8), 4 chains each, Z not kept and
    X'Z formed inside the draws: xtz_mode 3, the configuration's BF16 hi + lo
    split on the matrix cores (FP32 accumulation), or 2, FP32 FMA. Each
    analysis gets its lane split from the autotuner (xtz_mode 3 runs at four
    lanes per voxel); then `steps` sweeps of the whole batch are timed against
    the same sweeps run analysis by analysis. With `extras`, the same twelve
    analyses with FP32 X'Z and with Z stored, and a batch of small analyses."""
    from bsglmm_amd import Batch, Layout, Sampler, synth
    rows, samplers = [], []
    for k in range(12):
        d = synth.generate(f"c5_{k}")
        L = Layout(d["mask"], *d["grid"])
        S = Sampler(L, d["covar"], d["data"], seed=0xC5C5 + k, nchains=4, xtz_mode=xtz_mode, z_store=2)
        S.sweep(3)
        g, ms = S.autotune(reps=10)
        P, N = d["covar"].shape
        rows.append({"analysis": k, "V": L.totvox, "N": N, "P": P, "chains": 4, "lane_split_G": g,
                     "sweep_ms_by_G": {str(G): (None if m < 0 else round(float(m), 4)) for G, m in zip((1, 2, 4, 8), ms)}})
        samplers.append(S)
        del d
    draws = sum(4 * r["N"] * r["V"] for r in rows)
    seq, conc = baselines(samplers, steps, interleave=False)
    B = Batch(samplers)
    B.sweep(3)
    t0 = time.perf_counter()
    B.sweep(steps)
    bat = time.perf_counter() - t0
    B.sweep_graph(10, 10)
    t0 = time.perf_counter()
    B.sweep_graph(steps, 10)
    gra = time.perf_counter() - t0
    B.close()
    for S in samplers:
        S.close()
    out = {}
    if extras:
        out["fp32_xtz_batch"] = uneven_batch(steps, 2, False)
        out["small_analyses_batch"] = small_batch(max(steps, 100))
        out["precision_form_z_stored_batch"] = precision_batch(max(steps // 2, 10))
    form = ("FP32 state, Z not kept; X'Z as BF16 hi + lo products on MFMA (v_mfma_f32_16x16x16_bf16) inside the "
            "draws, FP32 accumulation" if xtz_mode == 3 else "FP32 state, Z not kept; X'Z on FP32 FMA inside the draws")
    return {**out, "form": form, "xtz_mode": xtz_mode, "analyses": rows, "chains_per_analysis": 4,
            "latent_draws_per_batch_sweep": draws,
            "batch_sweeps_per_s": steps / bat, "batched_draws_per_s": draws * steps / bat,
            "one_by_one_draws_per_s": draws * steps / seq, "batched_over_one_by_one": seq / bat,
            "concurrent_streams_draws_per_s": draws * steps / conc, "batched_over_concurrent_streams": conc / bat,
            "sum_of_tuned_sweep_ms": sum(min(v for v in r["sweep_ms_by_G"].values() if v is not None)
                                         for r in rows),
            "batch_sweep_ms": 1e3 * bat / steps, "graph_batch_sweep_ms": 1e3 * gra / steps}


def precision_batch(steps):
    """The same twelve analyses (4 chains each) in the configuration's stated
    form: Z stored (k_latent body), X'Z on the matrix cores with Z split into
    BF16 hi + lo (k_xtz_mfma body). Batched (one launch per phase and
    covariate count, each count on its own stream) against one at a time and
    against concurrent streams."""
    from bsglmm_amd import Batch, Layout, Sampler, synth
    samplers, draws = [], 0
    for k in range(12):
        d = synth.generate(f"c5_{k}")
        L = Layout(d["mask"], *d["grid"])
        samplers.append(Sampler(L, d["covar"], d["data"], seed=0xC5C5 + k, nchains=4, xtz_mode=1, z_store=0))
        P, N = d["covar"].shape
        draws += 4 * N * L.totvox
        del d
    for S in samplers:
        S.sweep(2)
        S.synchronize()
    seq, conc = baselines(samplers, steps, interleave=False)
    B = Batch(samplers)
    B.sweep(2)
    t0 = time.perf_counter()
    B.sweep(steps)
    bat = time.perf_counter() - t0
    B.close()
    for S in samplers:
        S.close()
    return {"form": "precision: Z stored, X'Z on MFMA with Z split into BF16 hi + lo, FP32 accumulation",
            "analyses": 12, "chains_per_analysis": 4, "timed_batch_sweeps": steps,
            "batch_sweeps_per_s": steps / bat, "batch_sweep_ms": 1e3 * bat / steps,
            "batched_draws_per_s": draws * steps / bat, "one_by_one_draws_per_s": draws * steps / seq,
            "batched_over_one_by_one": seq / bat, "concurrent_streams_draws_per_s": draws * steps / conc,
            "batched_over_concurrent_streams": conc / bat}


def baselines(samplers, steps, interleave):
    """Wall seconds of `steps` sweeps of every analysis without a batch:
    one analysis at a time (each synchronized before the next starts), and
    every analysis on its own stream with the launches issued together
    (sweep by sweep when `interleave`, else analysis by analysis), so the
    GPU overlaps them as it can."""
    t0 = time.perf_counter()
    for S in samplers:
        S.sweep(steps)
        S.synchronize()
    seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    if interleave:
        for _ in range(steps):
            for S in samplers:
                S.sweep(1)
    else:
        for S in samplers:
            S.sweep(steps)
    for S in samplers:
        S.synchronize()
    return seq, time.perf_counter() - t0


def small_batch(steps, count=24):
    """`count` small analyses (the 32^3 ellipsoid mask, V = 7,056; N from 20 to
    112, P from 2 to 8; one chain each), where a single analysis cannot fill
    the GPU: batched launches against the analyses swept one by one."""
    from bsglmm_amd import Batch, Layout, Sampler, synth
    samplers, draws = [], 0
    for k in range(count):
        n, p = 20 + 4 * k, 2 + k % 7
        d = synth.generate("c1", nsubs=n, ncovar=p)
        L = Layout(d["mask"], *d["grid"])
        samplers.append(Sampler(L, d["covar"], d["data"], seed=0x5A11 + k, xtz_mode=2, z_store=2))
        draws += n * L.totvox
    seq, conc = baselines(samplers, steps, interleave=True)
    B = Batch(samplers)
    B.sweep(3)
    t0 = time.perf_counter()
    B.sweep(steps)
    bat = time.perf_counter() - t0
    # the same batch sweeps replayed from HIP graphs of 50 (one host launch per group)
    B.sweep_graph(50, 50)
    t0 = time.perf_counter()
    B.sweep_graph(steps, 50)
    gra = time.perf_counter() - t0
    unified = B.one_launch_per_phase
    # and one launch per phase and covariate count, each P on its own stream
    B.set_launch_mode(2)
    B.sweep(3)
    t0 = time.perf_counter()
    B.sweep(steps)
    per_p = time.perf_counter() - t0
    B.close()
    for S in samplers:
        S.close()
    return {"analyses": count, "batch_sweeps_per_s": steps / bat, "batched_draws_per_s": draws * steps / bat,
            "one_by_one_draws_per_s": draws * steps / seq, "batched_over_one_by_one": seq / bat,
            "concurrent_streams_draws_per_s": draws * steps / conc, "batched_over_concurrent_streams": conc / bat,
            "graph_batch_sweeps_per_s": steps / gra, "graph_over_direct_batch": bat / gra,
            "graph_over_one_by_one": seq / gra, "one_launch_per_phase": unified,
            "per_p_launches_batch_sweeps_per_s": steps / per_p}


def load_profile(path):
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh)
    return {}


def main():
    a = parse()
    if a.cpu_leg:
        name, warm, timed, core = a.cpu_leg.split(":")
        print(json.dumps(cpu_leg(name, int(warm), int(timed), int(core)) or {}), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {a.gpus} is not supported "
              "(the sweep shards over voxel slabs by design but this build measures one GPU)",
              file=sys.stderr)
        sys.exit(2)
    # the reference CPU path on two pinned cores of their own, beside the GPU work
    legs = [] if a.no_cpu else [("c2", a.cpu_warmup, a.cpu_sweeps), ("c1", 0, a.cpu_c1_sweeps)]
    procs = start_cpu_legs(legs) if legs else []
    from bsglmm_amd import Layout, copy_bandwidth, sampler_rate, synth, valu_rate

    d = synth.generate(a.config)
    nx, ny, nz = d["grid"]
    L = Layout(d["mask"], nx, ny, nz)
    P, N = d["covar"].shape
    V = L.totvox

    # fused X'Z, Z not kept, overlapped; the probes (copy bandwidth, VALU
    # rates, the latent sampler's draw rate) run between building the context
    # and the warm-up sweeps, the sampler probe (~60 ms of the latent kernels'
    # own instruction mix) last, so that the sweeps start at the clock the
    # chip holds under that mix, as in a long run
    def settle():
        cb = copy_bandwidth(1 << 30, 100)
        tr = valu_rate("trans", 20)
        fm = valu_rate("fma", 100)
        sr = sampler_rate(100)
        return cb, fm, tr, sr

    head = time_mode(L, d, 2, 2, a.steps, a.warmup, settle=settle, graph=a.graph, persistent=a.persistent)
    if not head["overlapped"]:
        raise RuntimeError("the headline workload must run overlapped")
    (copy_gbs, copy_ms), (fma_rate, _), (trans_rate, _), (sampler_dps, _) = head["settle"]
    # overlap None: the context's default (on wherever supported)
    forms = {"z_free_serial": (2, 2, False), "fused_xtz_z_stored": (2, 0, True), "separate_xtz_fp32": (0, 0, True),
             "separate_xtz_mfma_bf16x2": (1, 0, True), "fused_xtz_mfma_bf16x2_z_free": (3, 2, None)}
    others = {k: time_mode(L, d, m, z, a.steps, a.warmup, o) for k, (m, z, o) in forms.items()}

    value = a.steps / head["wall"]
    sweep_bytes = N * V * (8 + 1 / 8) + V * (24 * P + 5)  # algorithmic bytes of one (unfused) sweep
    # the Z-free latent draws move y bits, b_j, WM_j in and m_j out: N V / 8 + 8 P V + 4 V
    vl_bytes = N * V / 8 + 8 * P * V + 4 * V
    # the two colour updates read m_j, the neighbours' b, the neighbour table
    # and WM_j and write b_j: 12 P V + 28 V; one sweep is two k_overlap launches
    ovl_bytes = (vl_bytes + 12 * P * V + 28 * V) / 2
    launch_s = head["lat_ms"] / 2 / 1e3  # the events bracket both launches of a sweep
    achieved = ovl_bytes / launch_s / 1e9
    prof = load_profile(a.profile)
    traffic = prof.get("bytes_per_launch")
    src = ("profile constant: " + prof.get("source", "") + "; divided by the launch time measured live here"
           if prof else None)
    limiter = None
    if prof.get("valu_wave_instructions_per_launch"):
        w = prof["valu_wave_instructions_per_launch"]
        # peak: 1024 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction; the
        # VALU probe of this run measures what one instruction kind reaches
        nominal = 1024 * 2.4e9 / 4
        limiter = {"kind": "valu_issue", "frac": w / (nominal * launch_s),
                   "wave_instructions_per_launch": w, "valu_per_draw": w * 2 * 64 / (N * V),
                   "peak": "1024 SIMDs x 2.4 GHz / 4 cycles per wave64 VALU instruction",
                   "peak_winst_per_s": nominal,
                   "probe_fma_winst_per_s": fma_rate, "frac_of_probe_fma_rate": w / (fma_rate * launch_s),
                   "probe_transcendental_winst_per_s": trans_rate,
                   "note": "the probe's all-FMA loop runs below the nominal rate (the chip's clock under it); this "
                           "kernel's integer/FP mix issues faster than that. SQ_INSTS_VALU counts a transcendental "
                           "(half rate: see the probe) as one instruction, so frac understates the issue slots used",
                   "source": src}

    def form_line(r, z_written, fused):
        lat_b = vl_bytes if not z_written else N * V * (4 + 1 / 8) + 4 * P * V + (4 * P * V if fused else 0)
        line = {"value": a.steps / r["wall"], "unit": "sweeps/s", "ms_per_step": 1e3 * r["wall"] / a.steps,
                "overlapped": r["overlapped"],
                "latent_kernel_ms": r["lat_ms"], "latent_gbs": lat_b / (r["lat_ms"] / 1e3) / 1e9,
                "kernel_ms_serial_breakdown": r["kms"],
                "effective_sweep_hbm_frac": sweep_bytes * (a.steps / r["wall"]) / 1e9 / HBM_PEAK_GBS}
        if not fused:
            line["xtz_gbs"] = (4 * N * V + 4 * P * V) / (r["kms"]["xtz"] / 1e3) / 1e9
        return line

    out = {
        "metric": METRIC, "value": value, "unit": "sweeps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * head["wall"] / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{a.config}: grid {nx}x{ny}x{nz} ellipsoidal-shell mask, plain single chain, "
                               f"FP32 state, X'Z on FP32 FMA fused into the latent draws (means x'b on FP32 MFMA), "
                               f"Z not kept, " + ("persistent sweep (one launch per call)" if a.persistent else
                                                 "overlapped sweep (two launches per sweep)"),
                   "V": V, "N": N, "P": P, "chains": 1},
        "latent_draws_per_s": value * N * V,
        "timed_path": (f"bsg_sweep_graph: the {a.steps} timed sweeps replayed from a HIP graph of "
                       f"{head['graph_group']} sweeps, captured before the warm-up (no capture in the timed region)"
                       if head["graph_group"] else "bsg_sweep: direct launches"),
        "direct_launch_sweeps_per_s": a.steps / head["direct_wall"] if head["direct_wall"] else value,
        "sweep_kernel_ms": head["lat_ms"],
        "k_overlap_avg_launch_ms": head["lat_ms"] / 2,
        "kernel_ms_serial_breakdown": head["kms"],
        "sweep_gpu_ms_with_latent_events": head["span_ms"],
        "sweep_roofline": {"kind": "effective: the unfused sweep's algorithmic bytes N V (8 + 1/8) + V (24 P + 5), "
                                   "Z written and read once, over the measured sweep time; this form never "
                                   "moves Z, so the figure credits bytes that are not moved",
                           "bound": "hbm", "algorithmic_bytes": sweep_bytes,
                           "achieved": sweep_bytes * value / 1e9, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": sweep_bytes * value / 1e9 / HBM_PEAK_GBS},
        "copy_kernel_gbs": copy_gbs,
        "valu_probe_winst_per_s": {"fma": fma_rate, "transcendental": trans_rate},
        "sampler_ceiling": {"draws_per_s": sampler_dps,
                            "what": "the latent sampler alone (Philox4x32-10 + inversion draw, no memory traffic), "
                                    "k_sampler_probe",
                            "sweep_latent_draws_frac": value * N * V / sampler_dps,
                            "latent_kernel_draws_frac": (N * V / 2) / (launch_s * sampler_dps)},
        "roofline": {"kernel": "k_overlap (one colour's latent draws + X'Z beside the other colour's update)",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": ovl_bytes, "traffic_source": src,
                     "hbm_counter_gbs": traffic / launch_s / 1e9 if traffic else None,
                     "limiter": limiter,
                     "note": "the kernel is bound by VALU issue (limiter), not by HBM: its algorithmic and "
                             "counted bytes are a few percent of the HBM roofline"},
        "forms": {k: form_line(r, forms[k][1] != 2, forms[k][0] >= 2) for k, r in others.items()},
        "cpu_baseline": None,
    }
    if a.config_lines:
        out["configs"] = {
            "c1": config_line("c1", [("z_free_overlapped", 2, 2, True, 0), ("z_free_overlapped_graph50", 2, 2, True, 50)],
                              a.c1_sweeps, 50),
            # C3's precision form with X'Z on the matrix cores: fused into the draws
            # (Z never in HBM), then with Z stored and read back; the FP32 fused form
            # beside them. At N = 1,000, P = 8 the serial order runs the Z-free forms
            # ~2 % faster than the overlapped one (the latent launch alone fills the
            # chip), so both are timed
            "c3": config_line("c3", [("precision_mfma_bf16x2_fused_xtz", 3, 2, False, 0),
                                     ("precision_mfma_bf16x2_fused_xtz_overlapped", 3, 2, True, 0),
                                     ("precision_mfma_bf16x2_xtz_z_stored", 1, 0, False, 0),
                                     ("z_free_fp32_fused_xtz", 2, 2, False, 0)], a.c3_sweeps, 10),
        }
    if a.batched_chains > 0:
        out["batched"] = batched_vs_plain(L, d, a.batched_chains, a.batched_iters, a.batched_burnin)
    if a.uneven_steps > 0:
        out["uneven_batch"] = uneven_batch(a.uneven_steps)
    if procs:
        out["cpu_baseline"] = finish_cpu_legs(procs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
