/*
 * bsglmm.h -- C ABI of the MI355X BSGLMM Gibbs sweep (libbsglmm.so).
 *
 * The reference has no plugin API: its sampler driver (bsglmm/mcmc.cpp)
 * calls C-linkage update functions that share state through process globals
 * (prototypes at mcmc.cpp:96-127). Each entry point below names the reference
 * function it replaces. Rules: every call returns an int status (BSG_OK or a
 * negative code; bsg_last_error() gives the message), nothing exits or
 * asserts; the caller owns host arrays, the context owns device memory; a
 * context issues all work on its own HIP stream; contexts are
 * thread-compatible, not thread-safe.
 *
 * Host arrays use the reference's layouts:
 *   mask        padded (NROW+2)(NCOL+2)(NDEPTH+2) bytes, raster index
 *               i + (NROW+2) j + (NROW+2)(NCOL+2) k        (main.cu:56, read_data.cpp:227-240)
 *   covar       float[P][N], covariate-major               (read_data.cpp:299-302)
 *   data        uint8[N][TOTVOX], y of subject i at mask rank  (read_data.cpp:334-395)
 *   WM          float[TOTVOX], NULL means 1.0 everywhere   (read_data.cpp:531-534)
 *   spatCoef    float[P][TOTVOXp], dilated-mask rank       (mcmc.cpp:227, covarGPU.cu:389-390)
 *   Z           float[N][TOTVOX]                            (mcmc.cpp:173-178)
 *   SpatCoefPrec float[P][P]                                (mcmc.cpp:150-154)
 */
#ifndef BSGLMM_H
#define BSGLMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSG_OK 0
#define BSG_ERR_ARG (-1)   /* bad argument or unsupported size */
#define BSG_ERR_HIP (-2)   /* HIP runtime error (no device, launch failure) */
#define BSG_ERR_NOMEM (-3) /* host or device allocation failed */
#define BSG_ERR_NUM (-4)   /* numerical failure reported by a kernel */

#define BSG_MAX_COVAR 20   /* P <= 20, as the reference (covarGPU.cu:65, 300-304). P <= 8 runs the
                              register-resident kernels in every form; 9 <= P <= 20 runs wide kernels
                              (P at run time, the P x P factor in LDS) in the Z-free fused form
                              (xtz_mode 2, z_store 2, latent_method 0), serial order, no batches */

typedef struct bsg_layout bsg_layout;
typedef struct bsg_ctx bsg_ctx;

/* Message of the last failing call on this thread. */
const char *bsg_last_error(void);

/* ---- layout: main.cu:189-394 (dilation, hostIdx, hostIdxSC, red/black
 * lists INDX[c].hostVox, neighbour counts INDX[c].hostNBRS via cnt_nbrs
 * main.cu:471-502; INDEX struct binCAR.h:12-19). Bit-exact with the
 * reference. ------------------------------------------------------------ */
int bsg_layout_create(const unsigned char *mask, int nrow, int ncol, int ndepth,
                      bsg_layout **out);
void bsg_layout_destroy(bsg_layout *layout);
/* counts = {TOTVOX, TOTVOXp, INDX[0].hostN, INDX[1].hostN} */
int bsg_layout_counts(const bsg_layout *layout, int counts[4]);
/* Copies the reference arrays out: hostIdx and hostIdxSC (NSIZE ints each),
 * INDX[c].hostVox (ints), INDX[c].hostNBRS (bytes). Any pointer may be NULL.
 * nbr6 (TOTVOX*6 ints): this library's colour-sorted neighbour table
 * (-x,+x,-y,+y,-z,+z; -1 outside the mask). */
int bsg_layout_export(const bsg_layout *layout, int *host_idx, int *host_idx_sc, int *vox0,
                      int *vox1, unsigned char *nbrs0, unsigned char *nbrs1, int *nbr6);

/* ---- sampler context ------------------------------------------------- */
typedef struct {
  uint64_t seed;        /* Philox key; the CLI derives it from seed.dat as main.cu:413 does */
  int nchains;          /* independent chains over the same data (>= 1) */
  int xtz_mode;         /* X'Z: 0 FP32 FMA kernel reading Z; 1 BF16 hi+lo split on MFMA,
                           FP32 accumulate; 2 FP32 FMA fused into the latent kernel;
                           3 BF16 hi+lo split on MFMA fused into the latent kernel (the
                           precision form without Z in HBM: z_store 2, latent_method 0,
                           P <= 8, no WM image, four lanes per voxel) */
  int df_offset;        /* Wishart df = TOTVOX + df_offset: 19 as covarGPU.cu:1104, P as covar.cpp:291 */
  float trunc_eps;      /* latent truncation: 1e-6 as covarGPU.cu:1414-1418, 0 as covar.cpp:317,320 */
  float pivot_min;      /* FP32 Cholesky pivot below this -> FP64 refactor of that voxel */
  int z_store;          /* Z in HBM: 0 stored, 1 stored with streaming (nontemporal) stores,
                           2 not kept (xtz_mode 2 and 3; the probit sweep needs Z only for X'Z) */
  int latent_method;    /* z_store 2 only: 0 every draw by inversion (the draws of z_store 0);
                           1 Box-Muller candidate kept inside the interval of y, exact
                           inversion redraw otherwise (an exact sampler with other draws) */
} bsg_opts;

void bsg_opts_default(bsg_opts *opts);

/* Allocates device state and uploads the data (read_data.cpp:321-332,
 * mcmc.cpp:136-253). Initial state as mcmc.cpp: beta = 0, Lambda = I, b = 0,
 * Z = 0, sweep index 0. The layout may be destroyed afterwards. */
int bsg_create(const bsg_layout *layout, int nsubs, int ncovar, const float *covar,
               const unsigned char *data, const float *wm, const bsg_opts *opts, bsg_ctx **out);
int bsg_destroy(bsg_ctx *ctx);

/* ---- the sweep (mcmc.cpp:311-394, GPU branch, probit) -------------------- */
/* nsweeps full sweeps: latent, X'Z, WM coefficient b (when a WM image was
 * given), coefficients colour 0 then 1, CAR precision. The CAR-precision draw of sweep t runs on a second internal
 * stream beside the latent draw of sweep t+1; the call returns with all work
 * ordered on the context's stream. */
int bsg_sweep(bsg_ctx *ctx, int nsweeps);
/* bsg_sweep through a HIP graph that holds sweeps_per_graph sweeps (captured
 * on first use, replayed for every full group; the remainder runs directly).
 * Same updates and random counters as bsg_sweep, so the same results.
 * Returns after the replays drained. */
int bsg_sweep_graph(bsg_ctx *ctx, int nsweeps, int sweeps_per_graph);
/* Captures and instantiates the graph of sweeps_per_graph sweeps that
 * bsg_sweep_graph would use, without running it (no sweep, the sweep index
 * unchanged), so that the first bsg_sweep_graph call only replays. */
int bsg_sweep_graph_prepare(bsg_ctx *ctx, int sweeps_per_graph);
/* updateZGPU (covarGPU.cu:1625-1669) */
int bsg_update_latent(bsg_ctx *ctx);
/* X'Z, the contraction fused in Spat_Coef_Probit (covarGPU.cu:327-352),
 * recomputed from Z (xtz_mode 2 and 3 already form it inside bsg_update_latent).
 * The b WM_j X'1 term of X'(Z - b WM) is applied by bsg_update_coef with the
 * b of the sweep. */
int bsg_compute_xtz(bsg_ctx *ctx);
/* updateBetaGPU (covarGPU.cu:686-724): the WM coefficient
 * b ~ N(S/W, 1/W), S = sum_j WM_j sum_i (Z_ij - x_i'b_j), W = N sum_j WM_j^2
 * (flat prior, as the GPU path). A no-op when the context has no WM image.
 * With z_store 2 it must follow bsg_update_latent of the same sweep. */
int bsg_update_wm_coef(bsg_ctx *ctx);
/* Spat_Coef_Probit for one colour (covarGPU.cu:547-560); reads X'Z from bsg_compute_xtz */
int bsg_update_coef(bsg_ctx *ctx, int colour);
/* updateSpatCoefPrecGPU (covarGPU.cu:1063-1112): FP64 pair reduction and a
 * device Wishart draw; advances the sweep index */
int bsg_update_car_prec(bsg_ctx *ctx);
/* waits for the context's stream; reports kernel-side numerical failures
 * (BSG_ERR_NUM: voxels whose posterior precision was not positive definite
 * in FP64 either kept their coefficients; the reference propagates NaNs
 * silently, covarGPU.cu:374), and BSG_ERR_HIP if an overlapped sweep's
 * colour-0 blocks waited past their bound for the CAR precision drawn beside
 * them. Each failure is reported once; the next sweeps start with it clear. */
int bsg_synchronize(bsg_ctx *ctx);
/* bsg_sweep with HIP events around every latent kernel: latent_ms receives
 * their summed durations, total_ms (optional) the whole span. Synchronous.
 * Overlapped sweeps (bsg_get_overlap): the events bracket each sweep's two
 * launches, which both draw latent variates. */
int bsg_sweep_timed(bsg_ctx *ctx, int nsweeps, double *latent_ms, double *total_ms);
/* nsweeps sweeps launched kernel by kernel with HIP events around each one;
 * kernel_ms[6] receives the summed durations of latent, xtz, coef0, coef1,
 * pairs, wishart (in a full sweep the colour-1 kernel also forms the pair
 * sums and draws the CAR precision, so pairs and wishart read 0). Returns
 * after the stream drained. */
int bsg_sweep_profiled(bsg_ctx *ctx, int nsweeps, double kernel_ms[6]);

/* ---- posterior accumulation on the device (mcmc.cpp:584-651) ----------- */
/* (Re)allocates and zeroes the accumulators of every chain: FP64 sums of
 * beta and beta^2, of the lesion probability Phi(beta_t + b WM) for the first
 * ntypes coefficients (compute_prbGPU, covarGPU.cu:1814-1824; Phi by
 * A&S 26.2.17 as covar.cpp:327-348) and of b WM; FP32 sums over nbatches
 * batches of batch_len draws for batch-means Monte-Carlo errors (draws past
 * nbatches * batch_len add to the sums but to no batch). */
int bsg_accum_reset(bsg_ctx *ctx, int ntypes, int nbatches, int batch_len);
/* the accumulators' shape as bsg_accum_reset (or a checkpoint read) set it;
 * zeros when none are allocated. Any pointer may be NULL. */
int bsg_accum_shape(bsg_ctx *ctx, int *ntypes, int *nbatches, int *batch_len);
/* adds the current draw of every chain */
int bsg_accum_add(bsg_ctx *ctx);
/* nsweeps sweeps, adding every save_every-th draw (mcmc.cpp:584: SAVE_ITER) */
int bsg_sweep_accumulate(bsg_ctx *ctx, int nsweeps, int save_every);
/* Posterior summaries of one chain in mask-rank order, any output may be NULL:
 * mean, var [P][TOTVOX] (variance with the n-1 denominator, mcmc.cpp:871-882),
 * prb [ntypes][TOTVOX], bwm_mean [TOTVOX], mcse [P+ntypes][TOTVOX] (batch
 * means over the completed batches, each of batch_len draws). ndraws receives
 * the number of draws. */
int bsg_accum_get(bsg_ctx *ctx, int chain, int *ndraws, double *mean, double *var, double *prb,
                  double *bwm_mean, double *mcse);

/* ---- per-save summaries of the current draw ----------------------------- */
/* alphaMean (covarGPU.cu:563-583): mean of each coefficient plane over the
 * mask, double[P] (FP64 sums). Synchronous. */
int bsg_coef_means(bsg_ctx *ctx, int chain, double *mean);
/* No counterpart in the reference: bsg_coef_means for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc). Same kernel, same block partials and the same order of sums, so
 * the same result bit for bit; the partials live in a fixed 64 KiB device
 * array per GPU instead of memory allocated on first use, and nothing is
 * allocated. Limit: 256 (P + 1) doubles <= 8192. Every call on one GPU shares
 * that array (each call overwrites all it reads): calls that overlap on one
 * GPU race. Synchronous. */
int bsg_coef_means_noalloc(bsg_ctx *ctx, int chain, double *mean);
/* compute_predictGPU (covarGPU.cu:1826-1957, MODEL 1): for every subject i and
 * subtype t < ntypes, ll[i*ntypes + t] = sum_j y_ij log(p + 1e-35) +
 * (1 - y_ij) log(1 - p + 1e-35) with p = Phi(b_t,j + b WM_j +
 * sum_{c >= ntypes} b_c,j x_c,i), Phi by A&S 26.2.17 in FP32 as
 * dProbSDNormGPU (covarGPU.cu:1672-1693); sums in FP64. The host forms DIC
 * and Qhat from it (covarGPU.cu:1927-1953, mcmc.cpp:719-786). Synchronous. */
int bsg_loglik(bsg_ctx *ctx, int chain, int ntypes, double *ll);

/* ---- state exchange in the reference layouts (chain index first) ------ */
int bsg_get_spat_coef(bsg_ctx *ctx, int chain, float *spat_coef);
int bsg_set_spat_coef(bsg_ctx *ctx, int chain, const float *spat_coef);
int bsg_get_z(bsg_ctx *ctx, int chain, float *z);
int bsg_get_xtz(bsg_ctx *ctx, int chain, float *m);           /* X'Z, float[P][TOTVOX] */
int bsg_set_xtz(bsg_ctx *ctx, int chain, const float *m);
int bsg_get_wm_coef(bsg_ctx *ctx, int chain, float *b);      /* WM coefficient b */
int bsg_set_wm_coef(bsg_ctx *ctx, int chain, float b);
int bsg_get_prec(bsg_ctx *ctx, int chain, float *prec);       /* float[P][P] */
int bsg_set_prec(bsg_ctx *ctx, int chain, const float *prec);
int bsg_get_pair_sums(bsg_ctx *ctx, int chain, double *t);    /* double[P][P], last update */
int bsg_get_sweep_index(bsg_ctx *ctx, int chain, uint32_t *sweep);
int bsg_set_sweep_index(bsg_ctx *ctx, uint32_t sweep);        /* all chains */
int bsg_get_fallback_count(bsg_ctx *ctx, uint32_t *count);     /* FP64 Cholesky refactors */

/* ---- checkpoint / resume: save_iter / restart_iter (mcmc.cpp:1265-1336,
 * RESTART main.cu:119, 152), exact. The file holds the whole chain state of
 * every chain: coefficients, X'Z, Lambda (FP32 and FP64), pair sums, b, the
 * sweep index, the Philox key, the lane split, the kernel-side counts and, if
 * allocated, the posterior accumulators. Z is not saved (the reference saves
 * it, mcmc.cpp:1288-1294): every draw of Z depends only on the current
 * coefficients, so the next sweep redraws it. A context created for the same
 * problem (mask, N, P, chains, Wishart df, truncation, pivot_min) and read
 * from the file continues the chains bit for bit: k sweeps, a checkpoint and
 * n - k sweeps in a new process equal n sweeps in one. Write only between
 * full sweeps; read replaces the key with the saved one. A checksum guards
 * the file. */
int bsg_checkpoint_write(bsg_ctx *ctx, const char *path);
int bsg_checkpoint_read(bsg_ctx *ctx, const char *path);

/* ---- tile shapes and batches of analyses (no counterpart in the reference:
 * it runs one analysis per process) ------------------------------------- */
/* Times whole Z-free sweeps (z_store 2) of this context at each lane split G
 * in {1, 2, 4, 8} lanes per voxel (voxels per workgroup = 256 / G, subject
 * tile = N / G), reps sweeps each from the same state, overlapped where the
 * split allows it (the split also decides that), restores the chain state
 * (coefficients, X'Z, Lambda, pair sums, b, sweep index), and keeps the
 * fastest split for later sweeps. best_g receives it; ms_per_g[4] (optional)
 * the mean sweep times. */
int bsg_autotune(bsg_ctx *ctx, int reps, int *best_g, double *ms_per_g);
/* Sets the lane split G of the Z-free latent kernel (1, 2, 4 or 8; 0 = chosen
 * from N), e.g. a split bsg_autotune found in an earlier process. Sweeps are
 * exact at every split; the split changes FP32 summation order only. */
int bsg_set_lane_split(bsg_ctx *ctx, int g);
int bsg_get_lane_split(bsg_ctx *ctx, int *g); /* the split the next sweep uses */
/* Overlapped sweeps (no counterpart in the reference, whose sweep runs its
 * kernels one after another, mcmc.cpp:311-394): full sweeps of a context
 * whose latent kernel runs at four lanes per voxel (inversion draws, no WM
 * image; any number of chains) run as two launches per sweep, each colour's
 * coefficient update beside the other colour's latent draws. Same variates,
 * same sums, same chains as with on = 0. On by default wherever supported
 * (P <= 8; at P = 7, 8 with fewer blocks per compute unit, and where the
 * latent launch alone fills the chip, as at N = 1,000, P = 8, the serial
 * order can be slightly faster). bsg_get_overlap tells whether the next
 * bsg_sweep* call runs overlapped. */
int bsg_set_overlap(bsg_ctx *ctx, int on);
int bsg_get_overlap(bsg_ctx *ctx, int *on);
/* Persistent sweeps (no counterpart in the reference, mcmc.cpp:311-394 runs
 * a sweep's kernels one after another): bsg_sweep's sweeps of an overlapped
 * one-chain context with P <= 6 and FP32 X'Z (xtz_mode 2) in one launch. A fixed grid of resident
 * blocks takes the sweeps' latent tiles, colour updates and CAR-precision
 * draws from a queue in an order that puts every item after what it reads,
 * so no launch boundary drains the chip between phases. Same variates, same
 * sums, same chains as the overlapped and the serial forms. Off by default;
 * bsg_sweep_accumulate and captured graphs keep the overlapped form.
 * bsg_get_persistent tells whether the next bsg_sweep runs persistent. The
 * lags (fractions of the resident block count, in (0, 64]) place each update
 * after its inputs in the queue; they change the speed, never the result. */
int bsg_set_persistent(bsg_ctx *ctx, int on);
int bsg_get_persistent(bsg_ctx *ctx, int *on);
int bsg_set_persistent_lags(bsg_ctx *ctx, double latent, double update, double draw);
/* Test hook of the overlapped sweep's bounded wait (no counterpart in the
 * reference): the colour-0 blocks give up after spin_limit polls (0: the
 * default 2^20) and the CAR-precision draw writes `tag` beside its entries
 * (0: the default 1, the only one the readers accept). With another tag every
 * wait runs out, the kernels still finish, and bsg_synchronize reports it. */
int bsg_test_overlap_wait(bsg_ctx *ctx, int spin_limit, int tag);
typedef struct bsg_batch bsg_batch;
/* Groups up to 32 contexts (independent analyses: own masks, N, P and chain
 * counts) so that each phase of a sweep is one kernel launch over all of
 * them: over every covariate count at once (four launches per batch sweep on
 * one stream, each block running the body of its member's P), or one launch
 * per phase and covariate count (each with its own register allocation, each
 * P on its own stream so that the launches of different P overlap). The
 * first is chosen for batches of fewer than 2^28 latent draws per sweep,
 * which cannot keep the chip busy between launches; bsg_batch_set_launch_mode
 * overrides the choice. Members must use a Z-free form (xtz_mode 2 or 3,
 * z_store 2: X'Z on FP32 FMA or as BF16 hi+lo products on MFMA inside the
 * draws) or the precision form with Z stored (xtz_mode 1, z_store 0,
 * latent_method 0: X'Z on MFMA from Z in HBM; a batch with such a member runs
 * one launch per phase and covariate count), have no WM image and share the latent method and sweep
 * index; each keeps its own device state and lane split, and must outlive the
 * batch. */
int bsg_batch_create(int n, bsg_ctx *const *members, bsg_batch **out);
int bsg_batch_destroy(bsg_batch *batch);
/* nsweeps sweeps of every member; returns when they are done. Waits for work
 * queued on the members' own streams first; BSG_ERR_ARG if the members'
 * sweep indices differ (a member was swept on its own since the last call). */
int bsg_batch_sweep(bsg_batch *batch, int nsweeps);
/* as bsg_sweep_accumulate for every member (bsg_accum_reset each first) */
int bsg_batch_sweep_accumulate(bsg_batch *batch, int nsweeps, int save_every);
/* bsg_batch_sweep through a HIP graph of sweeps_per_graph batch sweeps
 * (captured once, replayed for every full group; the remainder runs
 * directly): the host issues one graph launch per group instead of one
 * launch per phase and covariate count per sweep. Same results. */
int bsg_batch_sweep_graph(bsg_batch *batch, int nsweeps, int sweeps_per_graph);
/* mode 0: chosen from the draws per sweep; 1: one launch per phase over every
 * covariate count; 2: one launch per phase and covariate count. Same results
 * in every mode. bsg_batch_get_launch_mode tells whether the next sweep runs
 * one launch per phase (1) or per phase and covariate count (0). */
int bsg_batch_set_launch_mode(bsg_batch *batch, int mode);
int bsg_batch_get_launch_mode(bsg_batch *batch, int *unified);

/* ---- measurement utility (no counterpart in the reference) ------------- */
/* Times reps launches of a float4 stream-copy kernel over `bytes`; gbps is
 * (read + written bytes) / time. Calibrates bandwidth and counter figures.
 * The two buffers stay allocated (per device, for the process) so that no
 * free leaves the GPU idle between the probe and work timed after it. */
int bsg_measure_copy_bandwidth(size_t bytes, int reps, double *gbps, double *ms_per_copy);
/* kind 0: copy (read + write counted), 1: write-only stream, 2: read-only stream */
int bsg_measure_bandwidth(int kind, size_t bytes, int reps, double *gbps, double *ms_per_launch);
/* Times reps launches of a VALU-bound kernel (two rounds of eight resident
 * 256-thread blocks per compute unit, 8 independent chains per thread) and
 * returns the chip's wave64 vector-instruction rate: kind 0 FP32 FMA (full
 * rate), 1 the v_exp_f32 transcendental. The denominator of an issue-bound
 * kernel's roofline; run before timed work, it also leaves the GPU at the
 * clock it holds under such a kernel. */
int bsg_measure_valu_rate(int kind, int reps, double *winst_per_s, double *ms_per_launch);
/* Times reps launches of the latent sampler alone (Philox4x32-10 and the
 * truncated-normal inversion draw, no memory traffic; two rounds of eight
 * resident 256-thread blocks per compute unit) and returns its draws per
 * second: the ceiling of the latent kernels' draw rate on this device. */
int bsg_measure_sampler_rate(int reps, double *draws_per_s, double *ms_per_launch);

/* ---- host KISS, for seed.dat (randgen.cpp:9-23, 36-45; main.cu:54, 413) */
double bsg_kiss(unsigned long seed[3]);
/* devseed = runiform_long_n(ULONG_MAX, seed), as main.cu:413 */
uint64_t bsg_devseed(unsigned long seed[3]);

#ifdef __cplusplus
}
#endif
#endif
