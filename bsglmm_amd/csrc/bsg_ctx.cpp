// bsg_ctx.cpp -- C ABI of libbsglmm: context, data upload, sweep, state exchange.
#include <math.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>

#include "../../include/bsglmm.h"
#include "bsg_internal.h"

namespace bsg {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

}  // namespace bsg

using bsg::set_error;

struct bsg_layout {
  bsg::Layout L;
};

struct bsg_ctx {
  bsg::Layout L;
  bsg_opts opts{};
  int N = 0, P = 0, NS = 0, NQ = 0, NW = 0, C = 1;
  uint32_t sweep = 0;  // index of the next sweep (the Philox counter word), same for all chains
  hipStream_t stream = nullptr;   // all work; the caller-visible stream
  hipStream_t stream2 = nullptr;  // CAR-precision draw of sweep t beside the latent draw of t+1
  // colour set (bit c) whose leftover tiles' X'Z the last colour-tiled latent
  // launch left in partial sums, until a coefficient kernel (or bsg_get_xtz)
  // adds them
  int tail_pending = 0;
  // full sweeps run overlapped (k_overlap) where supported and asked for (by
  // default: at P <= 6 it shares the draws' register budget, at P = 7, 8 it
  // runs four blocks per compute unit so that the colour updates do not
  // spill); set_overlap
  bool overlap = true;
  bool overlap_ran = false;  // flags[2] (a Lambda wait past its bound) to check
  // pinned host copy of flags[0..3]: bsg_synchronize queues its read behind the
  // context's work and waits once for both
  uint32_t *flags_host = nullptr;
  bsg::DevProblem pb{};
  std::vector<void *> allocs;
  hipEvent_t ev[7] = {};
  // e_c0: colour-1 update done (stream); e_c1: Wishart draw done (stream2)
  hipEvent_t e_c0 = nullptr, e_c1 = nullptr;
  std::vector<hipEvent_t> lat_ev;  // timing pool for bsg_sweep_timed
  bsg::Accum acc;                  // posterior accumulators (bsg_accum_reset)
  std::vector<void *> acc_allocs;
  int acc_count = 0, acc_batch_len = 1;
  int bstat_parts = 0;             // WM-statistic partials written by the last latent launch
  double *cpart = nullptr;         // [chain][CM_BLOCKS][P] plane-mean partials
  double *llpart = nullptr, *llout = nullptr;  // log-likelihood partials / results (lazy)
  uint32_t *sweep_words = nullptr;  // [0] = 0 (direct launches), [1] = base of a captured group
  uint32_t sweep_stage = 0;         // host source of the base upload before graph replays
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_k = 0;
  bool graph_overlap = false;  // the form the graph was captured in
  // the persistent sweep (k_persist): asked for, its tagged words (allocated
  // on first use), the queue pattern and the launch epoch
  bool persist = false;
  bsg::PTags ptg{};
  size_t ptg_words = 0;       // tagged words from ptg.bt on (one allocation)
  uint32_t *ppat = nullptr;   // device copy of the pattern
  std::vector<uint32_t> ppat_host;
  int pwindows_extra = 0;     // windows past K that the pattern's offsets need
  int pblocks = 0;            // resident blocks launched
  double plag[3] = {1.0, 0.25, 0.1};  // latent, update, draw lags (fractions of pblocks)
  bool pplan_ok = false;

  void free_accum()
  {
    for (void *p : acc_allocs) hipFree(p);
    acc_allocs.clear();
    acc = bsg::Accum{};
    acc_count = 0;
  }

  void release()
  {
    if (stream) hipStreamSynchronize(stream);
    if (stream2) hipStreamSynchronize(stream2);
    if (graph_exec) hipGraphExecDestroy(graph_exec);
    if (graph) hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    free_accum();
    for (void *p : allocs) hipFree(p);
    allocs.clear();
    if (flags_host) hipHostFree(flags_host);
    flags_host = nullptr;
    for (auto &e : ev) {
      if (e) hipEventDestroy(e);
      e = nullptr;
    }
    for (auto &e : lat_ev) hipEventDestroy(e);
    lat_ev.clear();
    for (hipEvent_t *e : {&e_c0, &e_c1}) {
      if (*e) hipEventDestroy(*e);
      *e = nullptr;
    }
    if (stream2) hipStreamDestroy(stream2);
    if (stream) hipStreamDestroy(stream);
    stream = stream2 = nullptr;
  }
  ~bsg_ctx() { release(); }
};

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));               \
      return BSG_ERR_HIP;                                                         \
    }                                                                             \
  } while (0)

template <typename T>
static int dev_alloc(bsg_ctx *ctx, T **p, size_t count)
{
  void *ptr = nullptr;
  hipError_t e = hipMalloc(&ptr, count * sizeof(T) ? count * sizeof(T) : 16);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return BSG_ERR_NOMEM;
  }
  ctx->allocs.push_back(ptr);
  *p = static_cast<T *>(ptr);
  return BSG_OK;
}

// Posterior-accumulator batch of the next saved draw: draw k goes to batch
// k / batch_len while that is below nbatches; later draws add to the sums but
// to no batch (-1), so every completed batch holds exactly batch_len draws.
static int acc_batch_index(const bsg_ctx *ctx)
{
  const int b = ctx->acc_count / ctx->acc_batch_len;
  return b < ctx->acc.nbatches ? b : -1;
}

// A captured sweep graph holds the problem (lane split, wait hooks) by value:
// drop it whenever one of those changes, so the next graph call captures anew.
static void drop_graph(bsg_ctx *ctx)
{
  if (!ctx->graph_exec && !ctx->graph) return;
  hipStreamSynchronize(ctx->stream);
  if (ctx->graph_exec) hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) hipGraphDestroy(ctx->graph);
  ctx->graph_exec = nullptr;
  ctx->graph = nullptr;
  ctx->graph_k = 0;
}

extern "C" {

const char *bsg_last_error(void) { return bsg::g_last_error.c_str(); }

int bsg_layout_create(const unsigned char *mask, int nrow, int ncol, int ndepth, bsg_layout **out)
{
  if (!out) {
    set_error("bsg_layout_create: out is NULL");
    return BSG_ERR_ARG;
  }
  *out = nullptr;
  auto *lay = new (std::nothrow) bsg_layout;
  if (!lay) {
    set_error("bsg_layout_create: out of host memory");
    return BSG_ERR_NOMEM;
  }
  if (bsg::build_layout(mask, nrow, ncol, ndepth, lay->L) != 0) {
    delete lay;
    return BSG_ERR_ARG;
  }
  *out = lay;
  return BSG_OK;
}

void bsg_layout_destroy(bsg_layout *layout) { delete layout; }

int bsg_layout_counts(const bsg_layout *layout, int counts[4])
{
  if (!layout || !counts) {
    set_error("bsg_layout_counts: NULL argument");
    return BSG_ERR_ARG;
  }
  counts[0] = layout->L.totvox;
  counts[1] = layout->L.totvoxp;
  counts[2] = layout->L.n[0];
  counts[3] = layout->L.n[1];
  return BSG_OK;
}

int bsg_layout_export(const bsg_layout *layout, int *host_idx, int *host_idx_sc, int *vox0,
                      int *vox1, unsigned char *nbrs0, unsigned char *nbrs1, int *nbr6)
{
  if (!layout) {
    set_error("bsg_layout_export: NULL layout");
    return BSG_ERR_ARG;
  }
  const bsg::Layout &L = layout->L;
  // a colour list can be empty: copy only what exists (memcpy takes no null source)
  auto put = [](void *dst, const void *src, size_t bytes) {
    if (dst && bytes) memcpy(dst, src, bytes);
  };
  put(host_idx, L.host_idx.data(), sizeof(int) * L.nsize);
  put(host_idx_sc, L.host_idx_sc.data(), sizeof(int) * L.nsize);
  put(vox0, L.vox[0].data(), sizeof(int) * L.n[0]);
  put(vox1, L.vox[1].data(), sizeof(int) * L.n[1]);
  put(nbrs0, L.nbrs[0].data(), L.n[0]);
  put(nbrs1, L.nbrs[1].data(), L.n[1]);
  put(nbr6, L.nbr6.data(), sizeof(int) * L.nbr6.size());
  return BSG_OK;
}

void bsg_opts_default(bsg_opts *o)
{
  if (!o) return;
  o->seed = 0;
  o->nchains = 1;
  o->xtz_mode = 0;
  o->df_offset = 19;
  o->trunc_eps = 1e-6f;
  o->pivot_min = 1e-6f;
  o->z_store = 0;
  o->latent_method = 0;
}

int bsg_create(const bsg_layout *layout, int nsubs, int ncovar, const float *covar,
               const unsigned char *data, const float *wm, const bsg_opts *opts, bsg_ctx **out)
{
  if (!out) {
    set_error("bsg_create: out is NULL");
    return BSG_ERR_ARG;
  }
  *out = nullptr;
  if (!layout || !covar || !data) {
    set_error("bsg_create: NULL layout, covar or data");
    return BSG_ERR_ARG;
  }
  if (ncovar < 1 || ncovar > BSG_MAX_COVAR) {
    set_error("bsg_create: number of covariates must be in 1.." + std::to_string(BSG_MAX_COVAR));
    return BSG_ERR_ARG;
  }
  if (nsubs < 1) {
    set_error("bsg_create: no subjects");
    return BSG_ERR_ARG;
  }
  bsg_opts o;
  bsg_opts_default(&o);
  if (opts) o = *opts;
  if (o.nchains < 1 || o.nchains > 4096) {
    set_error("bsg_create: nchains must be in 1..4096");
    return BSG_ERR_ARG;
  }
  if (o.xtz_mode < 0 || o.xtz_mode > 3) {
    set_error("bsg_create: xtz_mode must be 0, 1, 2 or 3");
    return BSG_ERR_ARG;
  }
  if (o.z_store < 0 || o.z_store > 2 || (o.z_store == 2 && o.xtz_mode < 2)) {
    set_error("bsg_create: z_store must be 0, 1, or 2 with xtz_mode 2 or 3");
    return BSG_ERR_ARG;
  }
  if (o.xtz_mode == 3 && (o.z_store != 2 || o.latent_method != 0 || ncovar > bsg::MAX_P || wm)) {
    set_error("bsg_create: xtz_mode 3 (BF16 hi+lo X'Z on the matrix cores, fused into the draws) needs "
              "z_store 2, latent_method 0, at most " + std::to_string(bsg::MAX_P) + " covariates and no WM image");
    return BSG_ERR_ARG;
  }
  if (o.latent_method < 0 || o.latent_method > 1 || (o.latent_method == 1 && o.z_store != 2)) {
    set_error("bsg_create: latent_method must be 0, or 1 with z_store 2");
    return BSG_ERR_ARG;
  }
  if (ncovar > bsg::MAX_P && (o.z_store != 2 || o.xtz_mode != 2 || o.latent_method != 0)) {
    set_error("bsg_create: more than " + std::to_string(bsg::MAX_P) +
              " covariates run the Z-free fused form only (xtz_mode 2, z_store 2, latent_method 0)");
    return BSG_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("bsg_create: no HIP device");
    return BSG_ERR_HIP;
  }
  // the allocation-free entry points' scratch on this device (once per device)
  HIPCHK(bsg::scratch_init());

  std::unique_ptr<bsg_ctx> ctx(new (std::nothrow) bsg_ctx);
  if (!ctx) {
    set_error("bsg_create: out of host memory");
    return BSG_ERR_NOMEM;
  }
  ctx->L = layout->L;
  ctx->opts = o;
  ctx->overlap = true;
  const bsg::Layout &L = ctx->L;
  const int V = L.totvox, N = nsubs, P = ncovar, C = o.nchains;
  ctx->N = N;
  ctx->P = P;
  ctx->C = C;
  ctx->NS = (N + 3) & ~3;
  ctx->NQ = (N + 3) / 4;
  ctx->NW = (N + 31) / 32;
  const int NS = ctx->NS, NW = ctx->NW;
  const long long df = (long long)V + o.df_offset;
  if (df < P || df > (1LL << 30)) {
    set_error("bsg_create: Wishart degrees of freedom out of range");
    return BSG_ERR_ARG;
  }
  if ((long long)V * ctx->NQ >= (1LL << 32)) {
    set_error("bsg_create: TOTVOX * ceil(N/4) must be below 2^32 (Philox counter)");
    return BSG_ERR_ARG;
  }

  // ---- host-side preparation in sorted order ----
  // y indexes the two truncation intervals (covarGPU.cu:1414-1420), so only 0 and 1 are valid
  for (size_t k = 0; k < (size_t)N * V; k++)
    if (data[k] > 1) {
      set_error("bsg_create: lesion data must be 0 or 1 (value " + std::to_string((int)data[k]) +
                " at subject " + std::to_string(k / V) + ", mask rank " + std::to_string(k % V) + ")");
      return BSG_ERR_ARG;
    }
  std::vector<uint32_t> ybits((size_t)V * NW, 0u);
  for (int s = 0; s < V; s++) {
    const int rank = L.rank_of_sorted[s];
    for (int i = 0; i < N; i++)
      if (data[(size_t)i * V + rank] == 1) ybits[(size_t)s * NW + (i >> 5)] |= 1u << (i & 31);
  }
  std::vector<float> X((size_t)P * NS, 0.0f), xsum(P, 0.0f), xtx((size_t)P * P, 0.0f);
  for (int c = 0; c < P; c++)
    for (int i = 0; i < N; i++) X[(size_t)c * NS + i] = covar[(size_t)c * N + i];
  // X'X accumulated in float, subject-outer, as read_data.cpp:321-327
  for (int i = 0; i < N; i++)
    for (int a = 0; a < P; a++)
      for (int b = 0; b < P; b++) xtx[b + a * P] += covar[(size_t)a * N + i] * covar[(size_t)b * N + i];
  for (int c = 0; c < P; c++) {
    float acc = 0.0f;
    for (int i = 0; i < N; i++) acc += covar[(size_t)c * N + i];
    xsum[c] = acc;
  }
  std::vector<float> wms(V, 1.0f);
  if (wm)
    for (int s = 0; s < V; s++) wms[s] = wm[L.rank_of_sorted[s]];

  // ---- device state ----
  HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
  // pinned words for bsg_synchronize's read-back; without them it copies after the wait
  if (hipHostMalloc((void **)&ctx->flags_host, 4 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
    ctx->flags_host = nullptr;
    (void)hipGetLastError();
  }
  for (hipEvent_t *e : {&ctx->e_c0, &ctx->e_c1})
    HIPCHK(hipEventCreateWithFlags(e, hipEventDisableTiming));
  bsg::DevProblem &pb = ctx->pb;
  pb.V = V;
  pb.N = N;
  pb.NS = NS;
  pb.NQ = ctx->NQ;
  pb.NW = NW;
  pb.P = P;
  pb.nchains = C;
  pb.n0 = L.n[0];
  pb.vb = 0;
  pb.ve = V;
  pb.zfree = o.z_store == 2 ? 1 : 0;
  pb.xbx = o.xtz_mode == 3 ? 1 : 0;
  // b is drawn each sweep only when a WM image is given (mcmc.cpp:344-345, UPDATE_WM)
  pb.wm_update = wm ? 1 : 0;
  pb.latent_method = o.latent_method;
  if (P <= bsg::MAX_P && o.z_store == 2 && o.latent_method == 0 && bsg::latent_vl_lds_bytes(pb) > 65536) {
    set_error("bsg_create: z_store 2 stages X in LDS: 16 * 4 ceil(N/16) * P bytes must fit in 64 KiB "
              "(N * P <= 16384); use z_store 0 or 1");
    return BSG_ERR_ARG;
  }
  if (o.latent_method == 1 && (N > 32768 || bsg::latent_vl_lds_bytes(pb) > 65536)) {
    set_error("bsg_create: latent_method 1 needs N <= 32768 and ceil(N/4) * P * 16 bytes of X plus the "
              "fallback queues in 64 KiB of LDS");
    return BSG_ERR_ARG;
  }
  {
    double w2 = 0.0;
    for (int s = 0; s < V; s++) w2 += (double)wms[s] * (double)wms[s];
    pb.wm2 = (double)N * w2;
  }
  if (pb.wm_update && !(pb.wm2 > 0.0)) {
    set_error("bsg_create: the WM image is zero over the mask; b has no information");
    return BSG_ERR_ARG;
  }
  pb.key0 = (uint32_t)(o.seed & 0xffffffffu);
  pb.key1 = (uint32_t)(o.seed >> 32);
  pb.df = (int)df;
  pb.trunc_eps = o.trunc_eps;
  pb.pivot_min = o.pivot_min;
  const int NE = P * (P + 1) / 2;
  int rc;
  int *nbr6;
  uint32_t *yb;
  float *Xd, *xtxd, *xsumd, *wmd;
#define ALLOC(ptr, n)                                   \
  if ((rc = dev_alloc(ctx.get(), &(ptr), (n))) != BSG_OK) return rc;
  ALLOC(nbr6, (size_t)V * 6);
  ALLOC(yb, (size_t)V * NW);
  ALLOC(Xd, (size_t)P * NS);
  ALLOC(xtxd, (size_t)P * P);
  ALLOC(xsumd, (size_t)P);
  ALLOC(wmd, (size_t)V);
  float4 *xqd = nullptr;
  if (P > bsg::MAX_P) ALLOC(xqd, (size_t)ctx->NQ * P);
  uint16_t *xbfd = nullptr;  // BF16 hi/lo rows of X for the MFMA X'Z
  const int NSB = (NS + 7) & ~7;
  if (P <= bsg::MAX_P) ALLOC(xbfd, (size_t)2 * P * NSB);
  ALLOC(pb.Z, o.z_store == 2 ? (size_t)1 : (size_t)C * V * NS);
  ALLOC(pb.beta, (size_t)C * V * P);
  ALLOC(pb.M, (size_t)C * V * P);
  ALLOC(pb.lam, (size_t)C * P * P);
  ALLOC(pb.lam64, (size_t)C * P * P);
  ALLOC(pb.tsum, (size_t)C * P * P);
  pb.max_parts = std::max(bsg::PAIR_BLOCKS, (std::max(L.n[0], L.n[1]) + 255) / 256);
  ALLOC(pb.partials, (size_t)C * pb.max_parts * NE);

  ALLOC(pb.bwm, (size_t)C);
  ALLOC(pb.bstat, (size_t)C * bsg::BSTAT_SLOTS);

  // colour tiles of the Z-free latent launch at four lanes per voxel: each
  // colour's leftover 64-voxel tiles (T mod CUs) run as K pieces each when
  // that spreads them over the chip (k_latent_vl_mfm); K <= 8 (more, smaller
  // pieces measured no faster overlapped and slower serial: each piece pays a
  // block's set-up). One chain only: with several, chains are the grid's
  // second dimension, so one chain's pieces would be dispatched mid-launch,
  // where they balance nothing
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1) {
      set_error("bsg_create: cannot read the compute-unit count");
      return BSG_ERR_HIP;
    }
    pb.ncu = ncu;
    pb.ctile = (C == 1 && P <= bsg::MAX_P) ? 1 : 0;
    const int R = (N + 15) / 16;
    int mp = 0;
    for (int c = 0; c < 2; c++) {
      const int tiles = (L.n[c] + 63) / 64, tail = tiles % ncu;
      const int K = (tail > 0 && pb.ctile) ? std::min(std::min(ncu / tail, R), 8) : 0;
      pb.tl_left[c] = K > 1 ? tail : 0;
      pb.tl_k[c] = K > 1 ? K : 1;
      if (K > 1) mp += tail * K * 64 * P;
    }
    pb.mp_stride = mp;
    if (mp > 0) ALLOC(pb.mpart, (size_t)C * mp);
  }
  ALLOC(pb.lamtag, (size_t)C * P * P);
  ALLOC(pb.wbart, (size_t)C * P * P);

  ALLOC(pb.flags, (size_t)4);
  ALLOC(ctx->sweep_words, (size_t)2);
  pb.sweep_base = ctx->sweep_words;
#undef ALLOC
  pb.nbr6 = nbr6;
  pb.ybits = yb;
  pb.X = Xd;
  pb.Xq = xqd;
  pb.xbf = xbfd;
  pb.NSB = NSB;
  pb.xtx = xtxd;
  pb.xsum = xsumd;
  pb.wm = wmd;

  hipStream_t st = ctx->stream;
  HIPCHK(hipMemcpyAsync(nbr6, L.nbr6.data(), sizeof(int) * L.nbr6.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(yb, ybits.data(), sizeof(uint32_t) * ybits.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(Xd, X.data(), sizeof(float) * X.size(), hipMemcpyHostToDevice, st));
  if (xbfd) HIPCHK(bsg::launch_xbf(pb, xbfd, st));
  std::vector<float> xq;
  if (xqd) {  // [NQ][P] quads of X for the wide kernels, zero past N
    xq.assign((size_t)ctx->NQ * P * 4, 0.0f);
    for (int q = 0; q < ctx->NQ; q++)
      for (int c = 0; c < P; c++)
        for (int l = 0; l < 4; l++) xq[((size_t)q * P + c) * 4 + l] = X[(size_t)c * NS + 4 * q + l];
    HIPCHK(hipMemcpyAsync(xqd, xq.data(), sizeof(float) * xq.size(), hipMemcpyHostToDevice, st));
  }
  HIPCHK(hipMemcpyAsync(xtxd, xtx.data(), sizeof(float) * xtx.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(xsumd, xsum.data(), sizeof(float) * P, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(wmd, wms.data(), sizeof(float) * V, hipMemcpyHostToDevice, st));
  if (o.z_store != 2) HIPCHK(hipMemsetAsync(pb.Z, 0, sizeof(float) * (size_t)C * V * NS, st));
  HIPCHK(hipMemsetAsync(pb.beta, 0, sizeof(float) * (size_t)C * V * P, st));
  HIPCHK(hipMemsetAsync(pb.M, 0, sizeof(float) * (size_t)C * V * P, st));
  HIPCHK(hipMemsetAsync(pb.tsum, 0, sizeof(double) * (size_t)C * P * P, st));
  HIPCHK(hipMemsetAsync(pb.bwm, 0, sizeof(float) * C, st));

  HIPCHK(hipMemsetAsync(pb.flags, 0, sizeof(uint32_t) * 4, st));
  HIPCHK(hipMemsetAsync(pb.lamtag, 0, sizeof(unsigned long long) * (size_t)C * P * P, st));
  HIPCHK(hipMemsetAsync(ctx->sweep_words, 0, sizeof(uint32_t) * 2, st));

  std::vector<float> lam((size_t)C * P * P, 0.0f);
  std::vector<double> lam64((size_t)C * P * P, 0.0);
  for (int c = 0; c < C; c++)
    for (int a = 0; a < P; a++) {
      lam[(size_t)c * P * P + a * P + a] = 1.0f;  // Lambda = I (mcmc.cpp:150-154)
      lam64[(size_t)c * P * P + a * P + a] = 1.0;
    }
  HIPCHK(hipMemcpyAsync(pb.lam, lam.data(), sizeof(float) * lam.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(pb.lam64, lam64.data(), sizeof(double) * lam64.size(), hipMemcpyHostToDevice, st));
  for (auto &e : ctx->ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipStreamSynchronize(st));
  *out = ctx.release();
  return BSG_OK;
}

int bsg_destroy(bsg_ctx *ctx)
{
  delete ctx;
  return BSG_OK;
}

#define CTXCHK(ctx)                         \
  if (!(ctx)) {                             \
    set_error("NULL context");              \
    return BSG_ERR_ARG;                     \
  }

int bsg_update_latent(bsg_ctx *ctx)
{
  CTXCHK(ctx);
  HIPCHK(bsg::launch_latent(ctx->pb, ctx->opts.xtz_mode >= 2, ctx->opts.z_store, ctx->sweep, ctx->stream, 0,
                            ctx->L.totvox, &ctx->bstat_parts, 0, 0, &ctx->tail_pending));
  return BSG_OK;
}

static constexpr int CM_BLOCKS = 256;

int bsg_coef_means(bsg_ctx *ctx, int chain, double *mean)
{
  CTXCHK(ctx);
  if (chain < 0 || chain >= ctx->C || !mean) {
    set_error("bsg_coef_means: bad chain or NULL output");
    return BSG_ERR_ARG;
  }
  const int P = ctx->P;
  if (!ctx->cpart) {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (hipMalloc((void **)&ctx->cpart, sizeof(double) * ctx->C * CM_BLOCKS * P) != hipSuccess) {
      ctx->cpart = nullptr;
      set_error("bsg_coef_means: hipMalloc failed");
      return BSG_ERR_NOMEM;
    }
    ctx->allocs.push_back(ctx->cpart);
  }
  HIPCHK(bsg::launch_cmeans(ctx->pb, ctx->cpart, CM_BLOCKS, ctx->stream));
  std::vector<double> h((size_t)CM_BLOCKS * P);
  HIPCHK(hipMemcpyAsync(h.data(), ctx->cpart + (size_t)chain * CM_BLOCKS * P, sizeof(double) * h.size(),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int c = 0; c < P; c++) {
    double acc = 0.0;
    for (int b = 0; b < CM_BLOCKS; b++) acc += h[(size_t)b * P + c];
    mean[c] = acc / (double)ctx->L.totvox;
  }
  return BSG_OK;
}

int bsg_coef_means_noalloc(bsg_ctx *ctx, int chain, double *mean)
{
  CTXCHK(ctx);
  if (chain < 0 || chain >= ctx->C || !mean) {
    set_error("bsg_coef_means_noalloc: bad chain or NULL output");
    return BSG_ERR_ARG;
  }
  const int P = ctx->P, V = ctx->L.totvox;
  // partials [CM_BLOCKS][P], then the P folded sums
  if ((size_t)CM_BLOCKS * P + P > bsg::SCRATCH_DOUBLES) {
    set_error("bsg_coef_means_noalloc: " + std::to_string(CM_BLOCKS) + " x (P + 1) doubles exceed the " +
              std::to_string(bsg::SCRATCH_DOUBLES) + "-double scratch");
    return BSG_ERR_ARG;
  }
  double *scr = bsg::scratch_ptr();
  if (!scr) {
    set_error("bsg_coef_means_noalloc: no scratch on the current device (create the context with it current)");
    return BSG_ERR_HIP;
  }
  // the same kernel and block partials as bsg_coef_means, for this chain alone
  bsg::DevProblem one = ctx->pb;
  one.beta = ctx->pb.beta + (size_t)chain * V * P;
  one.nchains = 1;
  HIPCHK(bsg::launch_cmeans(one, scr, CM_BLOCKS, ctx->stream));
  HIPCHK(bsg::launch_cmeans_fold(scr, CM_BLOCKS, P, scr + (size_t)CM_BLOCKS * P, ctx->stream));
  HIPCHK(hipMemcpyAsync(mean, scr + (size_t)CM_BLOCKS * P, sizeof(double) * P, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int c = 0; c < P; c++) mean[c] /= (double)V;
  return BSG_OK;
}

int bsg_loglik(bsg_ctx *ctx, int chain, int ntypes, double *ll)
{
  CTXCHK(ctx);
  if (chain < 0 || chain >= ctx->C || !ll || ntypes < 1 || ntypes > ctx->P) {
    set_error("bsg_loglik: need a valid chain, 1 <= ntypes <= P and an output array");
    return BSG_ERR_ARG;
  }
  const int P = ctx->P, N = ctx->N, C = ctx->C;
  if (!ctx->llpart) {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const size_t tiles = (size_t)bsg::loglik_tiles(ctx->pb);
    if (hipMalloc((void **)&ctx->llpart, sizeof(double) * C * tiles * N * P) != hipSuccess ||
        hipMalloc((void **)&ctx->llout, sizeof(double) * C * N * P) != hipSuccess) {
      set_error("bsg_loglik: hipMalloc failed");
      return BSG_ERR_NOMEM;
    }
    ctx->allocs.push_back(ctx->llpart);
    ctx->allocs.push_back(ctx->llout);
  }
  HIPCHK(bsg::launch_loglik(ctx->pb, ntypes, ctx->llpart, ctx->llout, ctx->stream));
  HIPCHK(hipMemcpyAsync(ll, ctx->llout + (size_t)chain * N * ntypes, sizeof(double) * N * ntypes,
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BSG_OK;
}

int bsg_update_wm_coef(bsg_ctx *ctx)
{
  CTXCHK(ctx);
  if (!ctx->pb.wm_update) return BSG_OK;  // no WM image: b stays 0 (mcmc.cpp:344)
  if (ctx->bstat_parts == 0 && ctx->opts.z_store == 2) {
    set_error("bsg_update_wm_coef: Z is not kept; call bsg_update_latent first in this sweep");
    return BSG_ERR_ARG;
  }
  HIPCHK(bsg::launch_bwm(ctx->pb, ctx->bstat_parts, ctx->sweep, ctx->stream));
  ctx->bstat_parts = 0;
  return BSG_OK;
}

int bsg_compute_xtz(bsg_ctx *ctx)
{
  CTXCHK(ctx);
  if (ctx->opts.z_store == 2) {
    set_error("bsg_compute_xtz: this context does not keep Z (z_store 2); bsg_update_latent forms X'Z");
    return BSG_ERR_ARG;
  }
  HIPCHK(bsg::launch_xtz(ctx->pb, ctx->opts.xtz_mode, ctx->stream, 0, ctx->L.totvox));
  return BSG_OK;
}

int bsg_update_coef(bsg_ctx *ctx, int colour)
{
  CTXCHK(ctx);
  if (colour != 0 && colour != 1) {
    set_error("bsg_update_coef: colour must be 0 or 1");
    return BSG_ERR_ARG;
  }
  HIPCHK(bsg::launch_coef(ctx->pb, colour, false, ctx->sweep, ctx->stream, ctx->tail_pending));
  // colour 0 writes every leftover voxel's m_j, colour 1 its own
  ctx->tail_pending = colour == 0 ? 0 : (ctx->tail_pending & 1);
  return BSG_OK;
}

int bsg_update_car_prec(bsg_ctx *ctx)
{
  CTXCHK(ctx);
  HIPCHK(bsg::launch_pairs(ctx->pb, ctx->stream));
  HIPCHK(bsg::launch_wishart(ctx->pb, false, ctx->sweep, ctx->stream));
  ctx->sweep++;
  return BSG_OK;
}

static int accum_launch(bsg_ctx *ctx, hipStream_t st)
{
  if (!ctx->acc.sum) {
    set_error("posterior accumulation: call bsg_accum_reset first");
    return BSG_ERR_ARG;
  }
  HIPCHK(bsg::launch_accum(ctx->pb, ctx->acc, acc_batch_index(ctx), st));
  ctx->acc_count++;
  return BSG_OK;
}

// Full sweeps in the reference order (mcmc.cpp:311-394): latent draw (+ X'Z),
// colour-0 then colour-1 coefficients, CAR precision. The CAR-precision draw
// of sweep t (final pair-sum reduction and Wishart) only has to finish before
// the colour-0 update of sweep t+1, and nothing in the latent draw of t+1
// reads it. Where the Z-free kernel allows (latent_carries_wishart), block 0
// of the next latent launch draws it (w_inline), so a sweep is four launches
// on one stream with no event waits; the last sweep of a call draws it in a
// launch of its own. Otherwise it runs on stream2 beside the next latent draw,
// joined by events (wish_pending). Each call ends joined on `stream`.
// lat_ev (2 per sweep, optional) brackets each latent launch for timing.
// One sweep of problem pb, whose Philox sweep word is *pb.sweep_base + off.
// lat_ev2 (optional) brackets the latent launch.
static int enqueue_sweep(bsg_ctx *ctx, const bsg::DevProblem &pb, uint32_t off, hipEvent_t *lat_ev2,
                         bool &wish_pending, bool &w_inline, const bsg::Accum *ac = nullptr, int batch = 0)
{
  hipStream_t sa = ctx->stream, sb = ctx->stream2;
  const bool fused = ctx->opts.xtz_mode >= 2;
  const int V = ctx->L.totvox;
  if (lat_ev2) HIPCHK(hipEventRecord(lat_ev2[0], sa));
  int parts = 0, pieces = 0;
  HIPCHK(bsg::launch_latent(pb, fused, ctx->opts.z_store, off, sa, 0, V, &parts, w_inline ? 1 : 0, off - 1,
                            &pieces));
  ctx->tail_pending = 0;  // the colour kernels of this sweep add the pieces
  w_inline = false;
  if (lat_ev2) HIPCHK(hipEventRecord(lat_ev2[1], sa));
  if (!fused) HIPCHK(bsg::launch_xtz(pb, ctx->opts.xtz_mode, sa, 0, V));
  HIPCHK(bsg::launch_bwm(pb, parts, off, sa));  // WM coefficient (mcmc.cpp:344-345)
  if (wish_pending) HIPCHK(hipStreamWaitEvent(sa, ctx->e_c1, 0));  // Lambda of sweep t-1
  wish_pending = false;
  // a save sweep (ac): each colour update accumulates its voxels' new b;
  // colour 0 adds its own pieces and writes colour 1's m_j from theirs
  HIPCHK(bsg::launch_coef(pb, 0, false, off, sa, pieces, ac, batch));
  if (ctx->L.n[1] > 0) {
    HIPCHK(bsg::launch_coef(pb, 1, true, off, sa, 0, ac, batch));  // colour 1 + pair partials
    if (fused && ctx->opts.z_store == 2 && bsg::latent_carries_wishart(pb)) {
      w_inline = true;  // drawn by the next latent launch, or at the end of the call
    } else {
      HIPCHK(hipEventRecord(ctx->e_c0, sa));
      HIPCHK(hipStreamWaitEvent(sb, ctx->e_c0, 0));
      HIPCHK(bsg::launch_wishart(pb, true, off, sb));
      HIPCHK(hipEventRecord(ctx->e_c1, sb));
      wish_pending = true;
    }
  } else {
    HIPCHK(bsg::launch_pairs(pb, sa));
    HIPCHK(bsg::launch_wishart(pb, false, off, sa));
  }
  return BSG_OK;
}

// Full sweeps run overlapped (k_overlap) when the context allows it: Z-free
// inversion draws at four lanes per voxel, b fixed, both colours present and
// the Wishart draw small enough to carry (any number of chains).
static bool use_overlap(const bsg_ctx *ctx)
{
  return ctx->overlap && ctx->opts.z_store == 2 && bsg::overlap_supported(ctx->pb);
}

// k full sweeps of problem pb, Philox words *pb.sweep_base + off0 + i, in the
// overlapped form: colour 0's latent draws of the first sweep, then per sweep
// phase 0 (colour 1 drawn, colour 0 updated, Lambda of the sweep before drawn
// inside) and phase 1 (colour 1 updated, colour 0 of the next sweep drawn),
// then the last sweep's Lambda. Same variates and sums as enqueue_sweep.
// ev (optional, 2 per sweep) brackets each sweep's two launches. Save sweeps
// (accum_every > 0) accumulate inside the colour updates.
static int enqueue_overlap(bsg_ctx *ctx, const bsg::DevProblem &pb, uint32_t off0, int k, hipEvent_t *ev,
                           int accum_every)
{
  hipStream_t sa = ctx->stream;
  if (k <= 0) return BSG_OK;
  const int tail0 = pb.tl_k[0] > 1 ? 1 : 0, tail1 = pb.tl_k[1] > 1 ? 2 : 0;
  HIPCHK(bsg::launch_latent(pb, true, 2, off0, sa, 0, ctx->L.n[0]));
  for (int i = 0; i < k; i++) {
    const uint32_t off = off0 + (uint32_t)i;
    const bool save = accum_every > 0 && (i + 1) % accum_every == 0;
    const int batch = save ? acc_batch_index(ctx) : 0;
    const bsg::Accum *ac = save ? &ctx->acc : nullptr;
    if (ev) HIPCHK(hipEventRecord(ev[2 * i], sa));
    HIPCHK(bsg::launch_overlap(pb, 0, off, 1, i > 0 ? 1 : 0, off - 1, tail0, sa, ac, batch));
    HIPCHK(bsg::launch_overlap(pb, 1, off, i + 1 < k ? 1 : 0, 0, 0, tail1, sa, ac, batch));
    if (ev) HIPCHK(hipEventRecord(ev[2 * i + 1], sa));
    if (save) ctx->acc_count++;
  }
  HIPCHK(bsg::launch_wishart(pb, true, off0 + (uint32_t)k - 1, sa));
  ctx->tail_pending = 0;
  ctx->overlap_ran = true;
  return BSG_OK;
}

// ---------------------------------------------------------------------------
// The persistent sweep's queue (k_persist). One window holds one relative
// sweep's items: the latent items of each colour (whole tiles, then pieces),
// the 256-voxel updates of each colour, the Bartlett factor and the Lambda
// draw. Each item gets a time in queue positions: the latent items of colour
// 0, then of colour 1, spread over the window; an update `lat` positions
// after its own latent items and `upd` after the other colour's updates
// around it; the draw `upd` after the last update; colour 0's next updates
// `drw` after the draw; a latent item `upd` after the update of its voxels'
// last coefficients. lat, upd and drw are fractions of the resident block
// count (an item holds a block slot about that many positions), so that an
// item reaches a block about when its inputs are ready. The times are found
// as a fixed point over the periodic schedule (lags halved until one
// exists), then the window is sorted by time; an item of time t sits dk =
// floor(t / window) windows after its own sweep's. Correctness does not rest
// on the lags: check_persist_pattern verifies that every item comes after
// each item it waits for (the kernel's waits are the dependencies listed
// there), which is all that the queue needs to drain.
// ---------------------------------------------------------------------------
struct PItemRef {
  int type, idx;
};

struct PersistDeps {
  int nl[2] = {0, 0}, nu[2] = {0, 0};
  std::vector<int> lat_ublk[2];               // latent item -> update block of its voxels
  std::vector<std::vector<int>> lat_of[2];    // update block -> its latent items
  std::vector<std::pair<int, int>> nbr[2];    // update block -> range of the other colour's blocks around it
};

static void persist_deps(const bsg_ctx *ctx, const bsg::DevProblem &pb, PersistDeps &D)
{
  const int n0 = ctx->L.n[0], V = ctx->L.totvox;
  for (int c = 0; c < 2; c++) {
    const bsg::VlRange r = bsg::vl_range(pb, c);
    D.nl[c] = r.nmain + r.left * r.K;
    const int nv = c ? V - n0 : n0;
    D.nu[c] = (nv + 255) / 256;
    D.lat_ublk[c].resize(D.nl[c]);
    D.lat_of[c].assign(D.nu[c], {});
    for (int i = 0; i < D.nl[c]; i++) {
      const int tile = i < r.nmain ? i : r.nmain + (i - r.nmain) / r.K;
      const int u = std::min(tile / 4, D.nu[c] - 1);
      D.lat_ublk[c][i] = u;
      D.lat_of[c][u].push_back(i);
    }
    D.nbr[c].assign(D.nu[c], {INT32_MAX, -1});
    const int base = c ? n0 : 0, obase = c ? 0 : n0;
    for (int q = 0; q < nv; q++) {
      const int s = base + q, u = q / 256;
      for (int d = 0; d < 6; d++) {
        const int t = ctx->L.nbr6[(size_t)s * 6 + d];
        if (t < 0) continue;
        const int v = (t - obase) / 256;
        D.nbr[c][u].first = std::min(D.nbr[c][u].first, v);
        D.nbr[c][u].second = std::max(D.nbr[c][u].second, v);
      }
    }
  }
}

// every item after each item it waits for, over enough sweeps to see every
// offset; entries: dk << 28 | type << 24 | index
static bool check_persist_pattern(const PersistDeps &D, const std::vector<uint32_t> &pat, int dkmax)
{
  const int plen = (int)pat.size();
  // rank of each item in the window, and its offset
  std::vector<int> rl[2], ru[2], dl[2], du[2];
  int rw = -1, rb = -1, dw = 0, db = 0;
  for (int c = 0; c < 2; c++) {
    rl[c].assign(D.nl[c], -1);
    ru[c].assign(D.nu[c], -1);
    dl[c].assign(D.nl[c], 0);
    du[c].assign(D.nu[c], 0);
  }
  for (int j = 0; j < plen; j++) {
    const uint32_t e = pat[j];
    const int dk = (int)(e >> 28), type = (int)((e >> 24) & 15), idx = (int)(e & 0xFFFFFF);
    switch (type) {
    case bsg::PI_L0: case bsg::PI_L1: rl[type][idx] = j; dl[type][idx] = dk; break;
    case bsg::PI_U0: ru[0][idx] = j; du[0][idx] = dk; break;
    case bsg::PI_U1: ru[1][idx] = j; du[1][idx] = dk; break;
    case bsg::PI_W: rw = j; dw = dk; break;
    default: rb = j; db = dk; break;
    }
  }
  for (int c = 0; c < 2; c++) {
    for (int x : rl[c]) if (x < 0) return false;
    for (int x : ru[c]) if (x < 0) return false;
  }
  if (rw < 0 || rb < 0) return false;
  auto pos = [&](int k, int rank, int dk) { return (long long)(k + dk) * plen + rank; };
  const int Kv = dkmax + 4;
  for (int k = 0; k < Kv; k++) {
    for (int c = 0; c < 2; c++) {
      for (int i = 0; i < D.nl[c]; i++) {  // latent item: the update of k - 1 of its voxels
        const long long p = pos(k, rl[c][i], dl[c][i]);
        const int u = D.lat_ublk[c][i];
        if (k >= 1 && pos(k - 1, ru[c][u], du[c][u]) >= p) return false;
      }
      for (int u = 0; u < D.nu[c]; u++) {  // update: its latent items, the other colour around it, Lambda
        const long long p = pos(k, ru[c][u], du[c][u]);
        for (int i : D.lat_of[c][u])
          if (pos(k, rl[c][i], dl[c][i]) >= p) return false;
        const int ko = c ? k : k - 1, o = 1 - c;
        if (ko >= 0)
          for (int v = D.nbr[c][u].first; v <= D.nbr[c][u].second; v++)
            if (pos(ko, ru[o][v], du[o][v]) >= p) return false;
        if (k >= 1 && pos(k - 1, rw, dw) >= p) return false;
      }
    }
    const long long pw = pos(k, rw, dw);  // the draw: every update of k and the Bartlett factor
    for (int c = 0; c < 2; c++)
      for (int u = 0; u < D.nu[c]; u++)
        if (pos(k, ru[c][u], du[c][u]) >= pw) return false;
    if (pos(k, rb, db) >= pw) return false;
    if (k >= 2 && pos(k - 2, rw, dw) >= pos(k, rb, db)) return false;  // Bartlett: the draw of k - 2
  }
  return true;
}

static int build_persist_pattern(bsg_ctx *ctx, const bsg::DevProblem &pb)
{
  PersistDeps D;
  persist_deps(ctx, pb, D);
  const int T = D.nl[0] + D.nl[1] + D.nu[0] + D.nu[1] + 2;
  if (T >= (1 << 24) || D.nu[0] < 1 || D.nu[1] < 1) {
    set_error("persistent sweep: problem too large for the queue pattern");
    return BSG_ERR_ARG;
  }
  const double slots = std::max(1, ctx->pblocks);
  std::vector<double> tl[2], tu[2];
  double tw = 0.0, tb = 0.0;
  bool found = false;
  double scale = 1.0;
  for (int attempt = 0; attempt < 40 && !found; attempt++, scale *= 0.5) {
    const double DL = std::max(1.0, scale * ctx->plag[0] * slots), DU = std::max(1.0, scale * ctx->plag[1] * slots),
                 DW = std::max(1.0, scale * ctx->plag[2] * slots);
    std::vector<double> b[2];
    for (int c = 0; c < 2; c++) {
      b[c].resize(D.nl[c]);
      for (int i = 0; i < D.nl[c]; i++) b[c][i] = c * 0.5 * T + (i + 0.5) * 0.5 * T / D.nl[c];
      tl[c] = b[c];
      tu[c].assign(D.nu[c], 0.0);
    }
    tw = 0.0;
    tb = 0.0;
    bool stable = false;
    for (int it = 0; it < 200 && !stable; it++) {
      double change = 0.0;
      auto upd = [&](double &x, double v) {
        change = std::max(change, std::fabs(v - x));
        x = v;
      };
      for (int c = 0; c < 2; c++)
        for (int u = 0; u < D.nu[c]; u++) {
          double t = tw - T + DW;
          for (int i : D.lat_of[c][u]) t = std::max(t, tl[c][i] + DL);
          const int o = 1 - c;
          const double shift = c ? 0.0 : -(double)T;  // colour 0 waits for colour 1 of the sweep before
          for (int v = D.nbr[c][u].first; v <= D.nbr[c][u].second; v++) t = std::max(t, tu[o][v] + shift + DU);
          upd(tu[c][u], t);
        }
      double w = tb + 1.0;
      for (int c = 0; c < 2; c++)
        for (double x : tu[c]) w = std::max(w, x + DU);
      upd(tw, w);
      upd(tb, tw - T + 1.0);
      for (int c = 0; c < 2; c++)
        for (int i = 0; i < D.nl[c]; i++) upd(tl[c][i], std::max(b[c][i], tu[c][D.lat_ublk[c][i]] - T + DU));
      stable = change < 1e-9;
      if (tw > 64.0 * T) break;  // not periodic at this lag
    }
    found = stable;
  }
  if (!found) {
    set_error("persistent sweep: no periodic schedule");
    return BSG_ERR_ARG;
  }
  // window offsets and the sorted window
  struct E {
    double t;
    int dk;
    uint32_t code;
  };
  std::vector<E> es;
  es.reserve(T);
  double tmin = tw;
  for (int c = 0; c < 2; c++) {
    for (double x : tl[c]) tmin = std::min(tmin, x);
    for (double x : tu[c]) tmin = std::min(tmin, x);
  }
  tmin = std::min(tmin, tb);
  const double t0 = std::floor(tmin / T) * T;
  auto add = [&](double t, uint32_t type, int idx) {
    const double r = t - t0;
    const int dk = (int)std::floor(r / T);
    es.push_back({r - (double)dk * T, dk, (type << 24) | (uint32_t)idx});
  };
  for (int c = 0; c < 2; c++) {
    for (int i = 0; i < D.nl[c]; i++) add(tl[c][i], c ? bsg::PI_L1 : bsg::PI_L0, i);
    for (int u = 0; u < D.nu[c]; u++) add(tu[c][u], c ? bsg::PI_U1 : bsg::PI_U0, u);
  }
  add(tw, bsg::PI_W, 0);
  add(tb, bsg::PI_BT, 0);
  int dkmax = 0;
  for (const E &e : es) dkmax = std::max(dkmax, e.dk);
  if (dkmax > 15) {
    set_error("persistent sweep: schedule spans too many windows");
    return BSG_ERR_ARG;
  }
  std::stable_sort(es.begin(), es.end(), [](const E &a, const E &b) { return a.t < b.t; });
  std::vector<uint32_t> pat(es.size());
  for (size_t j = 0; j < es.size(); j++) pat[j] = ((uint32_t)es[j].dk << 28) | es[j].code;
  if (!check_persist_pattern(D, pat, dkmax)) {
    set_error("persistent sweep: the queue pattern puts an item before one it waits for");
    return BSG_ERR_ARG;
  }
  if (ctx->ppat && ctx->ppat_host.size() != pat.size()) {
    hipFree(ctx->ppat);
    auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), (void *)ctx->ppat);
    if (it != ctx->allocs.end()) ctx->allocs.erase(it);
    ctx->ppat = nullptr;
  }
  if (!ctx->ppat) {
    HIPCHK(hipMalloc((void **)&ctx->ppat, sizeof(uint32_t) * pat.size()));
    ctx->allocs.push_back(ctx->ppat);
  }
  HIPCHK(hipMemcpy(ctx->ppat, pat.data(), sizeof(uint32_t) * pat.size(), hipMemcpyHostToDevice));
  ctx->ppat_host = pat;
  ctx->pwindows_extra = dkmax;
  ctx->pplan_ok = true;
  return BSG_OK;
}

// the tagged words, the queue head and the pattern, on first use
static int persist_setup(bsg_ctx *ctx)
{
  const bsg::DevProblem &pb = ctx->pb;
  if (!ctx->ptg.bt) {
    const size_t V = ctx->L.totvox, P = ctx->P, NE = P * (P + 1) / 2;
    const size_t nu0 = (ctx->L.n[0] + 255) / 256, nu1 = (ctx->L.n[1] + 255) / 256;
    const size_t n_bt = 2 * V * P, n_mt = V * P, n_mpt = (size_t)std::max(pb.mp_stride, 1),
                 n_pt = nu1 * 2 * NE, n_lt = 2 * P * P, n_wb = 4 * P * P, n_u0 = nu0;
    if (!ctx->pblocks) ctx->pblocks = bsg::persist_max_blocks(pb);
    const size_t n_dg = 4 + 2 * (size_t)std::max(ctx->pblocks, 1) + 4 * 256;  // the wait diagnosis
    const size_t words = n_bt + n_mt + n_mpt + n_pt + n_lt + n_wb + n_u0 + n_dg;
    unsigned long long *w = nullptr;
    HIPCHK(hipMalloc((void **)&w, sizeof(unsigned long long) * words + 256));
    ctx->allocs.push_back(w);
    HIPCHK(hipMemset(w, 0, sizeof(unsigned long long) * words + 256));
    bsg::PTags &t = ctx->ptg;
    t.bt = w;
    t.mt = t.bt + n_bt;
    t.mpt = t.mt + n_mt;
    t.pt = t.mpt + n_mpt;
    t.lt = t.pt + n_pt;
    t.wb = t.lt + n_lt;
    t.u0d = t.wb + n_wb;
    ctx->pb.pdiag = t.u0d + n_u0;
    t.head = reinterpret_cast<uint32_t *>(t.u0d + n_u0 + n_dg);
    t.epoch = 0;
    ctx->ptg_words = words;
  }
  if (!ctx->pblocks) ctx->pblocks = bsg::persist_max_blocks(pb);
  if (ctx->pblocks < 1) {
    set_error("persistent sweep: the kernel cannot be resident on this device");
    return BSG_ERR_HIP;
  }
  if (!ctx->pplan_ok) return build_persist_pattern(ctx, pb);
  return BSG_OK;
}

static bool use_persist(const bsg_ctx *ctx)
{
  return ctx->persist && use_overlap(ctx) && bsg::persist_supported(ctx->pb);
}

// k sweeps of the persistent kernel (in launches of at most PERSIST_MAX_K);
// ev (optional, 2 per sweep): the first pair brackets each launch, the others
// are recorded back to back (so their sum is the launches' time)
static int enqueue_persist(bsg_ctx *ctx, int k, hipEvent_t *ev)
{
  int rc = persist_setup(ctx);
  if (rc != BSG_OK) return rc;
  hipStream_t sa = ctx->stream;
  int done = 0;
  const int plen = (int)ctx->ppat_host.size();
  const int kmax = std::min(bsg::PERSIST_MAX_K, (int)((1u << 31) / (uint32_t)plen) - ctx->pwindows_extra);
  if (kmax < 1) {
    set_error("persistent sweep: problem too large for one launch");
    return BSG_ERR_ARG;
  }
  while (done < k) {
    const int kk = std::min(k - done, kmax);
    bsg::PTags t = ctx->ptg;
    if (++ctx->ptg.epoch > (uint32_t)bsg::PERSIST_MAX_EPOCH) {
      // every tag of the earlier launches cleared before the epochs repeat
      HIPCHK(bsg::launch_zero64(ctx->ptg.bt, ctx->ptg_words, sa));
      ctx->ptg.epoch = 1;
    }
    t.epoch = ctx->ptg.epoch;
    t.pattern = ctx->ppat;
    t.plen = plen;
    t.K = kk;
    t.windows = kk + ctx->pwindows_extra;
    t.sweep0 = ctx->sweep + (uint32_t)done;
    t.tail = (ctx->pb.tl_k[0] > 1 ? 1 : 0) | (ctx->pb.tl_k[1] > 1 ? 2 : 0);
    if (ev) HIPCHK(hipEventRecord(ev[2 * done], sa));
    HIPCHK(bsg::launch_persist(ctx->pb, t, ctx->pblocks, sa));
    if (ev) {
      HIPCHK(hipEventRecord(ev[2 * done + 1], sa));
      for (int i = done + 1; i < done + kk; i++) {
        HIPCHK(hipEventRecord(ev[2 * i], sa));
        HIPCHK(hipEventRecord(ev[2 * i + 1], sa));
      }
    }
    done += kk;
  }
  ctx->tail_pending = 0;
  ctx->overlap_ran = true;
  return BSG_OK;
}

static int sweep_run(bsg_ctx *ctx, int nsweeps, hipEvent_t *lat_ev, int accum_every = 0)
{
  hipStream_t sa = ctx->stream;
  bool wish_pending = false, w_inline = false;
  if (accum_every > 0 && !ctx->acc.sum) {
    set_error("posterior accumulation: call bsg_accum_reset first");
    return BSG_ERR_ARG;
  }
  if (accum_every == 0 && use_persist(ctx)) {
    const int rc = enqueue_persist(ctx, nsweeps, lat_ev);
    if (rc == BSG_OK) ctx->sweep += (uint32_t)nsweeps;
    return rc;
  }
  if (use_overlap(ctx)) {
    const int rc = enqueue_overlap(ctx, ctx->pb, ctx->sweep, nsweeps, lat_ev, accum_every);
    if (rc == BSG_OK) ctx->sweep += (uint32_t)nsweeps;
    return rc;
  }
  for (int i = 0; i < nsweeps; i++) {
    // save sweeps accumulate inside the colour updates (mcmc.cpp:584-651 saves
    // after the sweep; each voxel's final b of the sweep is what it adds)
    const bool save = accum_every > 0 && (i + 1) % accum_every == 0;
    const int batch = save ? acc_batch_index(ctx) : 0;
    const int rc = enqueue_sweep(ctx, ctx->pb, ctx->sweep, lat_ev ? lat_ev + 2 * i : nullptr, wish_pending, w_inline,
                                 save ? &ctx->acc : nullptr, batch);
    if (rc != BSG_OK) return rc;
    ctx->sweep++;
    if (save) ctx->acc_count++;
  }
  if (w_inline) HIPCHK(bsg::launch_wishart(ctx->pb, true, ctx->sweep - 1, sa));  // the last sweep's draw
  if (wish_pending) HIPCHK(hipStreamWaitEvent(sa, ctx->e_c1, 0));  // join
  return BSG_OK;
}

// A HIP graph of k sweeps, captured once per k: the kernels read their sweep
// word as base + 0..k-1 from a device word that the graph's last node
// advances by k, so one executable graph serves every group of k sweeps.
static int build_graph(bsg_ctx *ctx, int k)
{
  if (ctx->graph_exec && ctx->graph_k == k && ctx->graph_overlap == use_overlap(ctx)) return BSG_OK;
  if (ctx->graph_exec) hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) hipGraphDestroy(ctx->graph);
  ctx->graph_exec = nullptr;
  ctx->graph = nullptr;
  hipStream_t sa = ctx->stream;
  bsg::DevProblem pg = ctx->pb;
  pg.sweep_base = ctx->sweep_words + 1;
  HIPCHK(hipStreamBeginCapture(sa, hipStreamCaptureModeThreadLocal));
  bool wish_pending = false, w_inline = false;
  int rc = BSG_OK;
  ctx->graph_overlap = use_overlap(ctx);
  if (ctx->graph_overlap) rc = enqueue_overlap(ctx, pg, 0, k, nullptr, 0);
  for (int i = 0; i < k && rc == BSG_OK && !ctx->graph_overlap; i++)
    rc = enqueue_sweep(ctx, pg, (uint32_t)i, nullptr, wish_pending, w_inline);
  if (rc == BSG_OK && w_inline && bsg::launch_wishart(pg, true, (uint32_t)(k - 1), sa) != hipSuccess) rc = BSG_ERR_HIP;
  if (rc == BSG_OK && wish_pending && hipStreamWaitEvent(sa, ctx->e_c1, 0) != hipSuccess) rc = BSG_ERR_HIP;
  if (rc == BSG_OK && bsg::launch_tick(ctx->sweep_words + 1, (uint32_t)k, sa) != hipSuccess) rc = BSG_ERR_HIP;
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(sa, &g);
  if (rc != BSG_OK || ec != hipSuccess) {
    if (g) hipGraphDestroy(g);
    if (rc == BSG_OK) set_error(std::string("graph capture: ") + hipGetErrorString(ec));
    return rc == BSG_OK ? BSG_ERR_HIP : rc;
  }
  ctx->graph = g;
  HIPCHK(hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
  ctx->graph_k = k;
  return BSG_OK;
}

int bsg_sweep_graph(bsg_ctx *ctx, int nsweeps, int sweeps_per_graph)
{
  CTXCHK(ctx);
  if (nsweeps < 0 || sweeps_per_graph < 1 || sweeps_per_graph > 256) {
    set_error("bsg_sweep_graph: need nsweeps >= 0 and 1 <= sweeps_per_graph <= 256");
    return BSG_ERR_ARG;
  }
  const int k = sweeps_per_graph, q = nsweeps / k;
  if (q > 0) {
    int rc = build_graph(ctx, k);
    if (rc != BSG_OK) return rc;
    ctx->sweep_stage = ctx->sweep;
    HIPCHK(hipMemcpyAsync(ctx->sweep_words + 1, &ctx->sweep_stage, sizeof(uint32_t), hipMemcpyHostToDevice,
                          ctx->stream));
    for (int j = 0; j < q; j++) HIPCHK(hipGraphLaunch(ctx->graph_exec, ctx->stream));
    ctx->sweep += (uint32_t)(q * k);
    // the staging word is read by the copy above: keep it stable until then
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  return sweep_run(ctx, nsweeps - q * k, nullptr);
}

int bsg_sweep_graph_prepare(bsg_ctx *ctx, int sweeps_per_graph)
{
  CTXCHK(ctx);
  if (sweeps_per_graph < 1 || sweeps_per_graph > 256) {
    set_error("bsg_sweep_graph_prepare: need 1 <= sweeps_per_graph <= 256");
    return BSG_ERR_ARG;
  }
  const int rc = build_graph(ctx, sweeps_per_graph);
  if (rc != BSG_OK) return rc;
  HIPCHK(hipGraphUpload(ctx->graph_exec, ctx->stream));  // its device-side setup, ahead of the first replay
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BSG_OK;
}

int bsg_sweep(bsg_ctx *ctx, int nsweeps)
{
  CTXCHK(ctx);
  if (nsweeps < 0) {
    set_error("bsg_sweep: negative sweep count");
    return BSG_ERR_ARG;
  }
  return sweep_run(ctx, nsweeps, nullptr);
}

int bsg_sweep_timed(bsg_ctx *ctx, int nsweeps, double *latent_ms, double *total_ms)
{
  CTXCHK(ctx);
  if (nsweeps < 1 || !latent_ms) {
    set_error("bsg_sweep_timed: bad argument");
    return BSG_ERR_ARG;
  }
  while ((int)ctx->lat_ev.size() < 2 * nsweeps + 2) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    ctx->lat_ev.push_back(e);
  }
  hipEvent_t *ev = ctx->lat_ev.data();
  HIPCHK(hipEventRecord(ev[2 * nsweeps], ctx->stream));
  int rc = sweep_run(ctx, nsweeps, ev);
  if (rc != BSG_OK) return rc;
  HIPCHK(hipEventRecord(ev[2 * nsweeps + 1], ctx->stream));
  HIPCHK(hipEventSynchronize(ev[2 * nsweeps + 1]));
  double acc = 0.0;
  for (int i = 0; i < nsweeps; i++) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
    acc += ms;
  }
  *latent_ms = acc;
  if (total_ms) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ev[2 * nsweeps], ev[2 * nsweeps + 1]));
    *total_ms = ms;
  }
  return BSG_OK;
}

int bsg_sweep_accumulate(bsg_ctx *ctx, int nsweeps, int save_every)
{
  CTXCHK(ctx);
  if (nsweeps < 0 || save_every < 1) {
    set_error("bsg_sweep_accumulate: bad sweep count or save interval");
    return BSG_ERR_ARG;
  }
  if (!ctx->acc.sum) {
    set_error("bsg_sweep_accumulate: call bsg_accum_reset first");
    return BSG_ERR_ARG;
  }
  return sweep_run(ctx, nsweeps, nullptr, save_every);
}

int bsg_accum_reset(bsg_ctx *ctx, int ntypes, int nbatches, int batch_len)
{
  CTXCHK(ctx);
  if (ntypes < 0 || ntypes > ctx->P || nbatches < 1 || nbatches > 1024 || batch_len < 1) {
    set_error("bsg_accum_reset: need 0 <= ntypes <= P, 1 <= nbatches <= 1024, batch_len >= 1");
    return BSG_ERR_ARG;
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->free_accum();
  const size_t CV = (size_t)ctx->C * ctx->L.totvox;
  const int P = ctx->P, K = P + ntypes;
  auto grab = [&](void **p, size_t bytes) -> int {
    if (hipMalloc(p, bytes ? bytes : 16) != hipSuccess) {
      set_error("bsg_accum_reset: hipMalloc failed");
      return BSG_ERR_NOMEM;
    }
    ctx->acc_allocs.push_back(*p);
    return hipMemsetAsync(*p, 0, bytes ? bytes : 16, ctx->stream) == hipSuccess ? BSG_OK : BSG_ERR_HIP;
  };
  bsg::Accum &a = ctx->acc;
  int rc;
  if ((rc = grab((void **)&a.sum, sizeof(double) * CV * P)) != BSG_OK ||
      (rc = grab((void **)&a.sq, sizeof(double) * CV * P)) != BSG_OK ||
      (rc = grab((void **)&a.prb, sizeof(double) * CV * ntypes)) != BSG_OK ||
      (rc = grab((void **)&a.bwm, sizeof(double) * CV)) != BSG_OK ||
      (rc = grab((void **)&a.batch, sizeof(float) * CV * nbatches * K)) != BSG_OK) {
    ctx->free_accum();
    return rc;
  }
  a.ntypes = ntypes;
  a.nbatches = nbatches;
  ctx->acc_batch_len = batch_len;
  ctx->acc_count = 0;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BSG_OK;
}

int bsg_accum_shape(bsg_ctx *ctx, int *ntypes, int *nbatches, int *batch_len)
{
  CTXCHK(ctx);
  const bool on = ctx->acc.sum != nullptr;
  if (ntypes) *ntypes = on ? ctx->acc.ntypes : 0;
  if (nbatches) *nbatches = on ? ctx->acc.nbatches : 0;
  if (batch_len) *batch_len = on ? ctx->acc_batch_len : 0;
  return BSG_OK;
}

int bsg_accum_add(bsg_ctx *ctx)
{
  CTXCHK(ctx);
  return accum_launch(ctx, ctx->stream);
}

int bsg_accum_get(bsg_ctx *ctx, int chain, int *ndraws, double *mean, double *var, double *prb,
                  double *bwm_mean, double *mcse)
{
  CTXCHK(ctx);
  if (chain < 0 || chain >= ctx->C || !ctx->acc.sum) {
    set_error("bsg_accum_get: bad chain or no accumulators");
    return BSG_ERR_ARG;
  }
  const int V = ctx->L.totvox, P = ctx->P, T = ctx->acc.ntypes, K = P + T, NB = ctx->acc.nbatches;
  const int n = ctx->acc_count;
  if (ndraws) *ndraws = n;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const size_t off = (size_t)chain * V;
  std::vector<double> sum((size_t)V * P), sq((size_t)V * P), pr((size_t)V * T), bw(V);
  HIPCHK(hipMemcpy(sum.data(), ctx->acc.sum + off * P, sizeof(double) * sum.size(), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(sq.data(), ctx->acc.sq + off * P, sizeof(double) * sq.size(), hipMemcpyDeviceToHost));
  if (T) HIPCHK(hipMemcpy(pr.data(), ctx->acc.prb + off * T, sizeof(double) * pr.size(), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(bw.data(), ctx->acc.bwm + off, sizeof(double) * V, hipMemcpyDeviceToHost));
  const double dn = n > 0 ? (double)n : 1.0;
  for (int s = 0; s < V; s++) {
    const int r = ctx->L.rank_of_sorted[s];
    for (int c = 0; c < P; c++) {
      const double m = sum[(size_t)s * P + c] / dn;
      if (mean) mean[(size_t)c * V + r] = m;
      // mcmc.cpp:878-881: (sum sq - n mean^2) / (n - 1)
      if (var) var[(size_t)c * V + r] = n > 1 ? (sq[(size_t)s * P + c] - dn * m * m) / (dn - 1.0) : 0.0;
    }
    for (int t = 0; t < T; t++)
      if (prb) prb[(size_t)t * V + r] = pr[(size_t)s * T + t] / dn;
    if (bwm_mean) bwm_mean[r] = bw[s] / dn;
  }
  if (mcse) {
    const int nb = std::min(n / ctx->acc_batch_len, NB);
    std::vector<float> bat((size_t)NB * V * K);
    HIPCHK(hipMemcpy(bat.data(), ctx->acc.batch + (size_t)chain * NB * V * K, sizeof(float) * bat.size(),
                     hipMemcpyDeviceToHost));
    for (int s = 0; s < V; s++) {
      const int r = ctx->L.rank_of_sorted[s];
      for (int k = 0; k < K; k++) {
        double m1 = 0.0, m2 = 0.0;
        for (int b = 0; b < nb; b++) {
          const double x = bat[((size_t)b * V + s) * K + k] / ctx->acc_batch_len;
          m1 += x;
          m2 += x * x;
        }
        double e = 0.0;
        if (nb > 1) {
          m1 /= nb;
          e = std::sqrt(std::max(0.0, (m2 - nb * m1 * m1) / (nb - 1)) / nb);
        }
        mcse[(size_t)k * V + r] = e;
      }
    }
  }
  return BSG_OK;
}

int bsg_synchronize(bsg_ctx *ctx)
{
  CTXCHK(ctx);
  // the second stream's work first; then the flags' read queued behind the
  // main stream's work, so that one wait covers both
  HIPCHK(hipStreamSynchronize(ctx->stream2));
  uint32_t flags[4];
  if (ctx->flags_host) {
    HIPCHK(hipMemcpyAsync(ctx->flags_host, ctx->pb.flags, sizeof(flags), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    memcpy(flags, ctx->flags_host, sizeof(flags));
  } else {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(flags, ctx->pb.flags, sizeof(flags), hipMemcpyDeviceToHost));
  }
  if (flags[2]) {
    // first: a wait that gave up leaves colour 0 with an incomplete precision,
    // which can also show as the failure below; both are reported here, once,
    // and the next sweeps start with the flags clear
    const uint32_t zero[3] = {0, 0, 0};
    HIPCHK(hipMemcpy(ctx->pb.flags + 1, zero, sizeof(zero), hipMemcpyHostToDevice));
    std::string where;
    if (flags[3] && ctx->pb.pdiag) {
      const bsg::PTags &t = ctx->ptg;
      const struct {
        const char *name;
        const void *p;
      } bufs[] = {{"coefficients", t.bt}, {"X'Z", t.mt}, {"piece partials", t.mpt}, {"pair partials", t.pt},
                  {"Lambda", t.lt}, {"Bartlett factor", t.wb}, {"colour-0 done words", t.u0d},
                  {"diagnosis", ctx->pb.pdiag}};
      auto name_of = [&](unsigned long long a, long long &off) {
        for (int b = 6; b >= 0; b--) {
          const unsigned long long base = (unsigned long long)(uintptr_t)bufs[b].p;
          if (base && a >= base && a < (unsigned long long)(uintptr_t)bufs[b + 1].p) {
            off = (long long)((a - base) / 8);
            return bufs[b].name;
          }
        }
        off = -1;
        return "k_overlap Lambda";
      };
      static const char *const items[] = {"L0", "L1", "U0", "U1", "W", "BT"};
      const size_t nb = ctx->pblocks, nrec = std::min<uint32_t>(flags[3], 16);  // the first 16 in the message
      std::vector<unsigned long long> all(4 + 2 * nb + 4 * 256);
      HIPCHK(hipMemcpy(all.data(), ctx->pb.pdiag, sizeof(unsigned long long) * all.size(), hipMemcpyDeviceToHost));
      const unsigned long long *rec = all.data() + 4 + 2 * nb;
      char buf[400];
      snprintf(buf, sizeof(buf), " (%u waits past their bound", flags[3]);
      where = buf;
      for (size_t r = 0; r < nrec; r++) {
        const unsigned long long *x = rec + 4 * r;
        long long off = 0;
        const char *nm = name_of(x[0], off);
        const uint32_t e = (uint32_t)x[3];
        snprintf(buf, sizeof(buf), "; %s word %lld wants %#llx has %#llx in %s(k=%u) #%u dk %u", nm, off,
                 (unsigned long long)x[1], x[2] >> 32, items[std::min<uint32_t>((e >> 24) & 15, 5)],
                 (uint32_t)(x[3] >> 32), e & 0xFFFFFFu, e >> 28);
        where += buf;
      }
      where += ")";
      HIPCHK(hipMemset(ctx->pb.pdiag + 4 + nb, 0, sizeof(unsigned long long) * (nb + 4 * 256)));
      HIPCHK(hipMemset(ctx->pb.flags + 3, 0, sizeof(uint32_t)));
    }
    set_error("overlapped or persistent sweep: a block waited past its bound for a value drawn in the same "
              "launch" + where);
    return BSG_ERR_HIP;
  }
  if (flags[1]) {
    // reported once: the next sweeps start with the count clear
    const uint32_t zero = 0;
    HIPCHK(hipMemcpy(ctx->pb.flags + 1, &zero, sizeof(zero), hipMemcpyHostToDevice));
    set_error("coefficient update: " + std::to_string(flags[1]) +
              " voxel(s) had a non positive definite posterior precision and kept their coefficients");
    return BSG_ERR_NUM;
  }
  return BSG_OK;
}

int bsg_sweep_profiled(bsg_ctx *ctx, int nsweeps, double kernel_ms[6])
{
  CTXCHK(ctx);
  if (!kernel_ms || nsweeps < 0) {
    set_error("bsg_sweep_profiled: bad argument");
    return BSG_ERR_ARG;
  }
  for (int k = 0; k < 6; k++) kernel_ms[k] = 0.0;
  hipStream_t st = ctx->stream;
  for (int i = 0; i < nsweeps; i++) {
    const bool fused = ctx->opts.xtz_mode >= 2;
    const uint32_t sw = ctx->sweep;
    HIPCHK(hipEventRecord(ctx->ev[0], st));
    int parts = 0, pieces = 0;
    HIPCHK(bsg::launch_latent(ctx->pb, fused, ctx->opts.z_store, sw, st, 0, ctx->L.totvox, &parts, 0, 0, &pieces));
    HIPCHK(hipEventRecord(ctx->ev[1], st));
    if (!fused) HIPCHK(bsg::launch_xtz(ctx->pb, ctx->opts.xtz_mode, st, 0, ctx->L.totvox));
    HIPCHK(bsg::launch_bwm(ctx->pb, parts, sw, st));
    HIPCHK(hipEventRecord(ctx->ev[2], st));
    ctx->tail_pending = 0;
    HIPCHK(bsg::launch_coef(ctx->pb, 0, false, sw, st, pieces));
    HIPCHK(hipEventRecord(ctx->ev[3], st));
    const bool one_pass = ctx->L.n[1] > 0;
    if (one_pass) HIPCHK(bsg::launch_coef(ctx->pb, 1, true, sw, st, 0));
    HIPCHK(hipEventRecord(ctx->ev[4], st));
    if (!one_pass) HIPCHK(bsg::launch_pairs(ctx->pb, st));
    HIPCHK(hipEventRecord(ctx->ev[5], st));
    HIPCHK(bsg::launch_wishart(ctx->pb, one_pass, sw, st));
    HIPCHK(hipEventRecord(ctx->ev[6], st));
    ctx->sweep = sw + 1;
    HIPCHK(hipEventSynchronize(ctx->ev[6]));
    for (int k = 0; k < 6; k++) {
      float ms = 0.0f;
      HIPCHK(hipEventElapsedTime(&ms, ctx->ev[k], ctx->ev[k + 1]));
      kernel_ms[k] += ms;
    }
  }
  return BSG_OK;
}

// ---- state exchange ----------------------------------------------------
#define CHAINCHK(ctx, chain)                   \
  CTXCHK(ctx);                                 \
  if ((chain) < 0 || (chain) >= (ctx)->C) {    \
    set_error("chain index out of range");     \
    return BSG_ERR_ARG;                        \
  }

int bsg_get_spat_coef(bsg_ctx *ctx, int chain, float *spat_coef)
{
  CHAINCHK(ctx, chain);
  const int V = ctx->L.totvox, VP = ctx->L.totvoxp, P = ctx->P;
  std::vector<float> b((size_t)V * P);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(b.data(), ctx->pb.beta + (size_t)chain * V * P, sizeof(float) * b.size(),
                   hipMemcpyDeviceToHost));
  memset(spat_coef, 0, sizeof(float) * (size_t)P * VP);
  for (int s = 0; s < V; s++)
    for (int c = 0; c < P; c++) spat_coef[(size_t)c * VP + ctx->L.sc_of_sorted[s]] = b[(size_t)s * P + c];
  return BSG_OK;
}

int bsg_set_spat_coef(bsg_ctx *ctx, int chain, const float *spat_coef)
{
  CHAINCHK(ctx, chain);
  const int V = ctx->L.totvox, VP = ctx->L.totvoxp, P = ctx->P;
  std::vector<float> b((size_t)V * P);
  for (int s = 0; s < V; s++)
    for (int c = 0; c < P; c++) b[(size_t)s * P + c] = spat_coef[(size_t)c * VP + ctx->L.sc_of_sorted[s]];
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(ctx->pb.beta + (size_t)chain * V * P, b.data(), sizeof(float) * b.size(),
                   hipMemcpyHostToDevice));
  return BSG_OK;
}

int bsg_get_z(bsg_ctx *ctx, int chain, float *z)
{
  CHAINCHK(ctx, chain);
  if (ctx->opts.z_store == 2) {
    set_error("bsg_get_z: this context does not keep Z (z_store 2)");
    return BSG_ERR_ARG;
  }
  const int V = ctx->L.totvox, N = ctx->N, NS = ctx->NS;
  std::vector<float> zz((size_t)V * NS);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(zz.data(), ctx->pb.Z + (size_t)chain * V * NS, sizeof(float) * zz.size(),
                   hipMemcpyDeviceToHost));
  for (int s = 0; s < V; s++) {
    const int rank = ctx->L.rank_of_sorted[s];
    for (int i = 0; i < N; i++) z[(size_t)i * V + rank] = zz[(size_t)s * NS + i];
  }
  return BSG_OK;
}

int bsg_get_xtz(bsg_ctx *ctx, int chain, float *m)
{
  CHAINCHK(ctx, chain);
  const int V = ctx->L.totvox, P = ctx->P;
  std::vector<float> mm((size_t)V * P);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(mm.data(), ctx->pb.M + (size_t)chain * V * P, sizeof(float) * mm.size(),
                   hipMemcpyDeviceToHost));
  for (int col = 0; col < 2; col++) {
    if (!((ctx->tail_pending >> col) & 1)) continue;
    // leftover tiles of this colour: add the pieces in piece order, as the
    // coefficient kernels do (the same FP32 sums)
    const bsg::VlRange r = bsg::vl_range(ctx->pb, col);
    const size_t per_tile = (size_t)r.K * 64 * P;
    std::vector<float> parts(per_tile * r.left);
    HIPCHK(hipMemcpy(parts.data(), ctx->pb.mpart + (size_t)chain * ctx->pb.mp_stride + r.moff,
                     sizeof(float) * parts.size(), hipMemcpyDeviceToHost));
    for (int s = r.base + r.nmain * 64; s < r.end; s++) {
      const int loc = s - r.base;
      const float *tp = parts.data() + (size_t)((loc >> 6) - r.nmain) * per_tile + (size_t)(loc & 63) * P;
      for (int c = 0; c < P; c++) {
        float sum = 0.0f;
        for (int kk = 0; kk < r.K; kk++) sum += tp[(size_t)kk * 64 * P + c];
        mm[(size_t)s * P + c] = sum;
      }
    }
  }
  for (int s = 0; s < V; s++)
    for (int c = 0; c < P; c++) m[(size_t)c * V + ctx->L.rank_of_sorted[s]] = mm[(size_t)s * P + c];
  return BSG_OK;
}

int bsg_set_xtz(bsg_ctx *ctx, int chain, const float *m)
{
  CHAINCHK(ctx, chain);
  const int V = ctx->L.totvox, P = ctx->P;
  std::vector<float> mm((size_t)V * P);
  for (int s = 0; s < V; s++)
    for (int c = 0; c < P; c++) mm[(size_t)s * P + c] = m[(size_t)c * V + ctx->L.rank_of_sorted[s]];
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(ctx->pb.M + (size_t)chain * V * P, mm.data(), sizeof(float) * mm.size(),
                   hipMemcpyHostToDevice));
  for (int col = 0; col < 2; col++) {
    if (!((ctx->tail_pending >> col) & 1)) continue;
    // this colour's leftover tiles still hold pieces: piece 0 = the new
    // value, the others 0, so the piece sum the coefficient kernels form is m
    const bsg::VlRange r = bsg::vl_range(ctx->pb, col);
    const size_t per_tile = (size_t)r.K * 64 * P;
    std::vector<float> parts(per_tile * r.left, 0.0f);
    for (int s = r.base + r.nmain * 64; s < r.end; s++) {
      const int loc = s - r.base;
      for (int c = 0; c < P; c++)
        parts[(size_t)((loc >> 6) - r.nmain) * per_tile + (size_t)(loc & 63) * P + c] = mm[(size_t)s * P + c];
    }
    HIPCHK(hipMemcpy(ctx->pb.mpart + (size_t)chain * ctx->pb.mp_stride + r.moff, parts.data(),
                     sizeof(float) * parts.size(), hipMemcpyHostToDevice));
  }
  return BSG_OK;
}

int bsg_get_wm_coef(bsg_ctx *ctx, int chain, float *b)
{
  CHAINCHK(ctx, chain);
  if (!b) {
    set_error("bsg_get_wm_coef: NULL output");
    return BSG_ERR_ARG;
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(b, ctx->pb.bwm + chain, sizeof(float), hipMemcpyDeviceToHost));
  return BSG_OK;
}

int bsg_set_wm_coef(bsg_ctx *ctx, int chain, float b)
{
  CHAINCHK(ctx, chain);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(ctx->pb.bwm + chain, &b, sizeof(float), hipMemcpyHostToDevice));
  return BSG_OK;
}

int bsg_get_prec(bsg_ctx *ctx, int chain, float *prec)
{
  CHAINCHK(ctx, chain);
  const int P = ctx->P;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(prec, ctx->pb.lam + (size_t)chain * P * P, sizeof(float) * P * P, hipMemcpyDeviceToHost));
  return BSG_OK;
}

int bsg_set_prec(bsg_ctx *ctx, int chain, const float *prec)
{
  CHAINCHK(ctx, chain);
  const int P = ctx->P;
  std::vector<double> d((size_t)P * P);
  for (int i = 0; i < P * P; i++) d[i] = prec[i];
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(ctx->pb.lam + (size_t)chain * P * P, prec, sizeof(float) * P * P, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->pb.lam64 + (size_t)chain * P * P, d.data(), sizeof(double) * P * P,
                   hipMemcpyHostToDevice));
  return BSG_OK;
}

int bsg_get_pair_sums(bsg_ctx *ctx, int chain, double *t)
{
  CHAINCHK(ctx, chain);
  const int P = ctx->P;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(t, ctx->pb.tsum + (size_t)chain * P * P, sizeof(double) * P * P, hipMemcpyDeviceToHost));
  return BSG_OK;
}

int bsg_get_sweep_index(bsg_ctx *ctx, int chain, uint32_t *sweep)
{
  CHAINCHK(ctx, chain);
  *sweep = ctx->sweep;
  return BSG_OK;
}

int bsg_set_sweep_index(bsg_ctx *ctx, uint32_t sweep)
{
  CTXCHK(ctx);
  ctx->sweep = sweep;
  return BSG_OK;
}

// ---- checkpoint / resume (the reference's save_iter / restart_iter,
// mcmc.cpp:1265-1336, made exact) ------------------------------------------
namespace {

constexpr char CKPT_MAGIC[8] = {'B', 'S', 'G', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t CKPT_VERSION = 1;

struct CkptHeader {
  char magic[8];
  uint32_t version;
  int32_t V, N, P, C, n0, lat_g, df, acc_present;
  uint32_t key0, key1, sweep, pad0;
  float trunc_eps, pivot_min;
  uint32_t flags[4];
  int32_t ntypes, nbatches, batch_len, acc_count;
  uint64_t layout_hash;
};

// FNV-1a, 64 bit
uint64_t fnv(uint64_t h, const void *p, size_t n)
{
  const unsigned char *c = static_cast<const unsigned char *>(p);
  for (size_t i = 0; i < n; i++) {
    h ^= c[i];
    h *= 1099511628211ull;
  }
  return h;
}

uint64_t layout_hash(const bsg::Layout &L)
{
  uint64_t h = 1469598103934665603ull;
  h = fnv(h, L.rank_of_sorted.data(), sizeof(int) * L.rank_of_sorted.size());
  h = fnv(h, L.nbr6.data(), sizeof(int) * L.nbr6.size());
  return h;
}

// the device arrays of the chain state, in file order
struct Part {
  void *dev;
  size_t bytes;
};

std::vector<Part> state_parts(bsg_ctx *ctx, bool with_acc)
{
  const size_t C = ctx->C, V = ctx->L.totvox, P = ctx->P;
  const bsg::DevProblem &pb = ctx->pb;
  std::vector<Part> v = {{pb.beta, 4 * C * V * P}, {pb.M, 4 * C * V * P},     {pb.lam, 4 * C * P * P},
                         {pb.lam64, 8 * C * P * P}, {pb.tsum, 8 * C * P * P}, {pb.bwm, 4 * C}};
  if (with_acc) {
    const bsg::Accum &a = ctx->acc;
    const size_t T = a.ntypes, NB = a.nbatches;
    v.push_back({a.sum, 8 * C * V * P});
    v.push_back({a.sq, 8 * C * V * P});
    if (T) v.push_back({a.prb, 8 * C * V * T});
    v.push_back({a.bwm, 8 * C * V});
    v.push_back({a.batch, 4 * C * NB * V * (P + T)});
  }
  return v;
}

}  // namespace

int bsg_checkpoint_write(bsg_ctx *ctx, const char *path)
{
  CTXCHK(ctx);
  if (!path) {
    set_error("bsg_checkpoint_write: NULL path");
    return BSG_ERR_ARG;
  }
  if (ctx->tail_pending) {
    set_error("bsg_checkpoint_write: a latent step's X'Z is still in pieces; finish the sweep first");
    return BSG_ERR_ARG;
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream2));
  CkptHeader h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, CKPT_MAGIC, 8);
  h.version = CKPT_VERSION;
  h.V = ctx->L.totvox;
  h.N = ctx->N;
  h.P = ctx->P;
  h.C = ctx->C;
  h.n0 = ctx->L.n[0];
  h.lat_g = ctx->pb.lat_g;
  h.df = ctx->pb.df;
  h.key0 = ctx->pb.key0;
  h.key1 = ctx->pb.key1;
  h.sweep = ctx->sweep;
  h.trunc_eps = ctx->pb.trunc_eps;
  h.pivot_min = ctx->pb.pivot_min;
  HIPCHK(hipMemcpy(h.flags, ctx->pb.flags, sizeof(h.flags), hipMemcpyDeviceToHost));
  h.acc_present = ctx->acc.sum ? 1 : 0;
  if (h.acc_present) {
    h.ntypes = ctx->acc.ntypes;
    h.nbatches = ctx->acc.nbatches;
    h.batch_len = ctx->acc_batch_len;
    h.acc_count = ctx->acc_count;
  }
  h.layout_hash = layout_hash(ctx->L);
  FILE *f = fopen(path, "wb");
  if (!f) {
    set_error(std::string("bsg_checkpoint_write: cannot open ") + path);
    return BSG_ERR_ARG;
  }
  uint64_t sum = fnv(1469598103934665603ull, &h, sizeof(h));
  bool ok = fwrite(&h, sizeof(h), 1, f) == 1;
  std::vector<char> buf;
  for (const Part &q : state_parts(ctx, h.acc_present != 0)) {
    if (!ok) break;
    buf.resize(q.bytes);
    if (hipMemcpy(buf.data(), q.dev, q.bytes, hipMemcpyDeviceToHost) != hipSuccess) {
      fclose(f);
      set_error("bsg_checkpoint_write: device read failed");
      return BSG_ERR_HIP;
    }
    sum = fnv(sum, buf.data(), q.bytes);
    ok = fwrite(buf.data(), 1, q.bytes, f) == q.bytes;
  }
  ok = ok && fwrite(&sum, sizeof(sum), 1, f) == 1;
  ok = (fclose(f) == 0) && ok;
  if (!ok) {
    set_error(std::string("bsg_checkpoint_write: write to ") + path + " failed");
    return BSG_ERR_ARG;
  }
  return BSG_OK;
}

int bsg_checkpoint_read(bsg_ctx *ctx, const char *path)
{
  CTXCHK(ctx);
  if (!path) {
    set_error("bsg_checkpoint_read: NULL path");
    return BSG_ERR_ARG;
  }
  FILE *f = fopen(path, "rb");
  if (!f) {
    set_error(std::string("bsg_checkpoint_read: cannot open ") + path);
    return BSG_ERR_ARG;
  }
  CkptHeader h;
  auto fail = [&](const std::string &msg) {
    fclose(f);
    set_error("bsg_checkpoint_read: " + msg);
    return BSG_ERR_ARG;
  };
  if (fread(&h, sizeof(h), 1, f) != 1 || memcmp(h.magic, CKPT_MAGIC, 8) != 0 || h.version != CKPT_VERSION)
    return fail(std::string(path) + " is not a checkpoint of this version");
  if (h.V != ctx->L.totvox || h.N != ctx->N || h.P != ctx->P || h.C != ctx->C || h.n0 != ctx->L.n[0] ||
      h.layout_hash != layout_hash(ctx->L))
    return fail("the checkpoint is of another problem (mask, N, P or chain count differ)");
  if (h.df != ctx->pb.df || h.trunc_eps != ctx->pb.trunc_eps || h.pivot_min != ctx->pb.pivot_min)
    return fail("the checkpoint was written under other sampler options (Wishart df, truncation or pivot_min)");
  if (h.acc_present && (h.ntypes < 0 || h.ntypes > h.P || h.nbatches < 1 || h.nbatches > 1024 || h.batch_len < 1 ||
                        h.acc_count < 0))
    return fail("bad accumulator header");
  if (h.lat_g != 0 && h.lat_g != 1 && h.lat_g != 2 && h.lat_g != 4 && h.lat_g != 8) return fail("bad lane split");
  // read everything, check the sum, then touch the context
  uint64_t sum = fnv(1469598103934665603ull, &h, sizeof(h));
  std::vector<std::vector<char>> data;
  {
    const size_t C = h.C, V = h.V, P = h.P, T = h.ntypes, NB = h.nbatches;
    std::vector<size_t> sizes = {4 * C * V * P, 4 * C * V * P, 4 * C * P * P, 8 * C * P * P, 8 * C * P * P, 4 * C};
    if (h.acc_present) {
      sizes.insert(sizes.end(), {8 * C * V * P, 8 * C * V * P});
      if (T) sizes.push_back(8 * C * V * T);
      sizes.insert(sizes.end(), {8 * C * V, 4 * C * NB * V * (P + T)});
    }
    for (size_t n : sizes) {
      data.emplace_back(n);
      if (fread(data.back().data(), 1, n, f) != n) return fail("the file is truncated");
      sum = fnv(sum, data.back().data(), n);
    }
  }
  uint64_t stored = 0;
  if (fread(&stored, sizeof(stored), 1, f) != 1 || stored != sum) return fail("checksum mismatch");
  fclose(f);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream2));
  if (h.acc_present) {
    const int rc = bsg_accum_reset(ctx, h.ntypes, h.nbatches, h.batch_len);
    if (rc != BSG_OK) return rc;
  } else {
    ctx->free_accum();
  }
  const std::vector<Part> parts = state_parts(ctx, h.acc_present != 0);
  for (size_t k = 0; k < parts.size(); k++)
    HIPCHK(hipMemcpy(parts[k].dev, data[k].data(), parts[k].bytes, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->pb.flags, h.flags, sizeof(h.flags), hipMemcpyHostToDevice));
  drop_graph(ctx);  // a captured graph holds the key and the split by value
  ctx->pb.key0 = h.key0;
  ctx->pb.key1 = h.key1;
  ctx->pb.lat_g = h.lat_g;
  ctx->sweep = h.sweep;
  ctx->acc_count = h.acc_present ? h.acc_count : 0;
  ctx->tail_pending = 0;
  ctx->bstat_parts = 0;
  return BSG_OK;
}

int bsg_get_fallback_count(bsg_ctx *ctx, uint32_t *count)
{
  CTXCHK(ctx);
  uint32_t flags[4];
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(flags, ctx->pb.flags, sizeof(flags), hipMemcpyDeviceToHost));
  *count = flags[0];
  return BSG_OK;
}

// The probe's buffers stay allocated for the process (per device, grown on
// demand): freeing them leaves the GPU idle for milliseconds, and work timed
// right after a probe would start below the GPU's working clock.
static std::mutex g_probe_mu;
static void *g_probe_buf[2][bsg::MAX_DEVICES];
static size_t g_probe_bytes[bsg::MAX_DEVICES];

int bsg_measure_bandwidth(int kind, size_t bytes, int reps, double *gbps, double *ms_per_launch)
{
  if (!gbps || bytes < 16 || reps < 1 || kind < 0 || kind > 2) {
    set_error("bsg_measure_bandwidth: bad argument");
    return BSG_ERR_ARG;
  }
  bytes &= ~(size_t)15;
  std::lock_guard<std::mutex> lock(g_probe_mu);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= bsg::MAX_DEVICES) {
    set_error("bsg_measure_bandwidth: no device");
    return BSG_ERR_HIP;
  }
  if (g_probe_bytes[dev] < bytes) {
    for (int b = 0; b < 2; b++) {
      if (g_probe_buf[b][dev]) hipFree(g_probe_buf[b][dev]);
      g_probe_buf[b][dev] = nullptr;
    }
    g_probe_bytes[dev] = 0;
  }
  void *src = g_probe_buf[0][dev], *dst = g_probe_buf[1][dev];
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = BSG_OK;
  float ms = 0.0f;
  if (!src && (hipMalloc(&src, bytes) != hipSuccess || hipMalloc(&dst, bytes) != hipSuccess)) {
    set_error("bsg_measure_bandwidth: hipMalloc failed");
    if (src) hipFree(src);
    src = dst = nullptr;
    rc = BSG_ERR_NOMEM;
  } else if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
             hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
             hipMemsetAsync(src, 0, bytes, st) != hipSuccess ||
             bsg::launch_probe(kind, src, dst, bytes, st) != hipSuccess ||
             hipEventRecord(e0, st) != hipSuccess) {
    set_error("bsg_measure_bandwidth: HIP failure");
    rc = BSG_ERR_HIP;
  } else {
    for (int i = 0; i < reps && rc == BSG_OK; i++)
      if (bsg::launch_probe(kind, src, dst, bytes, st) != hipSuccess) rc = BSG_ERR_HIP;
    if (rc == BSG_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = BSG_ERR_HIP;
    if (rc != BSG_OK) set_error("bsg_measure_bandwidth: HIP failure");
  }
  if (rc == BSG_OK) {
    const double moved = (kind == 0 ? 2.0 : 1.0) * (double)bytes;
    *gbps = moved * reps / (ms * 1e-3) / 1e9;
    if (ms_per_launch) *ms_per_launch = ms / reps;
  }
  if (st) hipStreamSynchronize(st);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (st) hipStreamDestroy(st);
  if (src && !g_probe_buf[0][dev]) {
    g_probe_buf[0][dev] = src;
    g_probe_buf[1][dev] = dst;
    g_probe_bytes[dev] = bytes;
  }
  return rc;
}

int bsg_measure_copy_bandwidth(size_t bytes, int reps, double *gbps, double *ms_per_copy)
{
  return bsg_measure_bandwidth(0, bytes, reps, gbps, ms_per_copy);
}

// the VALU probe's output (one float per block; written only if a chain hits
// an impossible value), per device, kept for the process like the buffers above
static float *g_valu_out[bsg::MAX_DEVICES];
static constexpr int VALU_BLOCKS_MAX = 1 << 16;

int bsg_measure_valu_rate(int kind, int reps, double *winst_per_s, double *ms_per_launch)
{
  if (!winst_per_s || reps < 1 || kind < 0 || kind > 1) {
    set_error("bsg_measure_valu_rate: bad argument");
    return BSG_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_probe_mu);
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= bsg::MAX_DEVICES ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) {
    set_error("bsg_measure_valu_rate: no device");
    return BSG_ERR_HIP;
  }
  // two rounds of eight resident 256-thread blocks per compute unit; about
  // 1 ms per launch (FMA) on 256 compute units
  const int blocks = std::min(16 * cus, VALU_BLOCKS_MAX), iters = 4096;
  if (!g_valu_out[dev] && hipMalloc((void **)&g_valu_out[dev], sizeof(float) * VALU_BLOCKS_MAX) != hipSuccess) {
    g_valu_out[dev] = nullptr;
    set_error("bsg_measure_valu_rate: hipMalloc failed");
    return BSG_ERR_NOMEM;
  }
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = BSG_OK;
  float ms = 0.0f;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess || bsg::launch_valu(kind, g_valu_out[dev], blocks, iters, st) != hipSuccess ||
      hipEventRecord(e0, st) != hipSuccess) {
    rc = BSG_ERR_HIP;
  } else {
    for (int i = 0; i < reps && rc == BSG_OK; i++)
      if (bsg::launch_valu(kind, g_valu_out[dev], blocks, iters, st) != hipSuccess) rc = BSG_ERR_HIP;
    if (rc == BSG_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.0f))
      rc = BSG_ERR_HIP;
  }
  if (rc != BSG_OK) set_error("bsg_measure_valu_rate: HIP failure");
  if (rc == BSG_OK) {
    const double winst = (double)blocks * 4.0 * iters * bsg::VALU_CHAINS_PER_ROUND;
    *winst_per_s = winst * reps / (ms * 1e-3);
    if (ms_per_launch) *ms_per_launch = ms / reps;
  }
  if (st) hipStreamSynchronize(st);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (st) hipStreamDestroy(st);
  return rc;
}

int bsg_measure_sampler_rate(int reps, double *draws_per_s, double *ms_per_launch)
{
  if (!draws_per_s || reps < 1) {
    set_error("bsg_measure_sampler_rate: bad argument");
    return BSG_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_probe_mu);
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= bsg::MAX_DEVICES ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) {
    set_error("bsg_measure_sampler_rate: no device");
    return BSG_ERR_HIP;
  }
  // two rounds of eight resident 256-thread blocks per compute unit; about
  // 1 ms per launch on 256 compute units
  const int blocks = std::min(16 * cus, VALU_BLOCKS_MAX), rounds = 128;
  if (!g_valu_out[dev] && hipMalloc((void **)&g_valu_out[dev], sizeof(float) * VALU_BLOCKS_MAX) != hipSuccess) {
    g_valu_out[dev] = nullptr;
    set_error("bsg_measure_sampler_rate: hipMalloc failed");
    return BSG_ERR_NOMEM;
  }
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = BSG_OK;
  float ms = 0.0f;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess || bsg::launch_sampler_probe(g_valu_out[dev], blocks, rounds, st) != hipSuccess ||
      hipEventRecord(e0, st) != hipSuccess) {
    rc = BSG_ERR_HIP;
  } else {
    for (int i = 0; i < reps && rc == BSG_OK; i++)
      if (bsg::launch_sampler_probe(g_valu_out[dev], blocks, rounds, st) != hipSuccess) rc = BSG_ERR_HIP;
    if (rc == BSG_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.0f))
      rc = BSG_ERR_HIP;
  }
  if (rc != BSG_OK) set_error("bsg_measure_sampler_rate: HIP failure");
  if (rc == BSG_OK) {
    const double draws = (double)blocks * 256.0 * rounds * 4.0;
    *draws_per_s = draws * reps / (ms * 1e-3);
    if (ms_per_launch) *ms_per_launch = ms / reps;
  }
  if (st) hipStreamSynchronize(st);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (st) hipStreamDestroy(st);
  return rc;
}

// ---- host KISS (randgen.cpp:9-23, 36-45; main.cu:54) -------------------
double bsg_kiss(unsigned long seed[3])
{
  static const double M = exp(-64 * log(2.0));
  seed[1] = seed[1] ^ (seed[1] << 17);
  seed[2] = (seed[2] ^ (seed[2] << 18)) & 0x7FFFFFFF;
  seed[0] = 69069 * seed[0] + 23606797;
  seed[1] ^= (seed[1] >> 15);
  seed[2] ^= (seed[2] >> 13);
  return (seed[0] + seed[1] + seed[2]) * M;
}

uint64_t bsg_devseed(unsigned long seed[3])
{
  const unsigned long n = ~0UL;
  unsigned long i = (unsigned long)floorl(n * bsg_kiss(seed));
  return (uint64_t)(i == n ? i - 1 : i);
}

}  // extern "C"

int bsg_set_lane_split(bsg_ctx *ctx, int g)
{
  CTXCHK(ctx);
  if (g != 0 && g != 1 && g != 2 && g != 4 && g != 8) {
    set_error("bsg_set_lane_split: G must be 0, 1, 2, 4 or 8");
    return BSG_ERR_ARG;
  }
  if (ctx->pb.xbx && g != 0 && g != 4) {
    set_error("bsg_set_lane_split: xtz_mode 3 forms X'Z on the matrix cores at G = 4 only");
    return BSG_ERR_ARG;
  }
  if (ctx->pb.lat_g != g) drop_graph(ctx);
  ctx->pb.lat_g = g;
  return BSG_OK;
}

int bsg_get_lane_split(bsg_ctx *ctx, int *g)
{
  CTXCHK(ctx);
  if (!g) {
    set_error("bsg_get_lane_split: NULL output");
    return BSG_ERR_ARG;
  }
  *g = ctx->pb.lat_g > 0 ? ctx->pb.lat_g : bsg::latent_vl_groups(ctx->pb);
  return BSG_OK;
}

int bsg_test_overlap_wait(bsg_ctx *ctx, int spin_limit, int tag)
{
  CTXCHK(ctx);
  if (spin_limit < 0 || tag < 0) {
    set_error("bsg_test_overlap_wait: negative argument");
    return BSG_ERR_ARG;
  }
  drop_graph(ctx);
  ctx->pb.lam_spin = spin_limit;
  ctx->pb.lam_tag = tag;
  return BSG_OK;
}

int bsg_set_persistent(bsg_ctx *ctx, int on)
{
  CTXCHK(ctx);
  ctx->persist = on != 0;
  return BSG_OK;
}

int bsg_get_persistent(bsg_ctx *ctx, int *on)
{
  CTXCHK(ctx);
  if (!on) {
    set_error("bsg_get_persistent: null argument");
    return BSG_ERR_ARG;
  }
  *on = use_persist(ctx) ? 1 : 0;
  return BSG_OK;
}

int bsg_set_persistent_lags(bsg_ctx *ctx, double latent, double update, double draw)
{
  CTXCHK(ctx);
  if (!(latent > 0.0 && update > 0.0 && draw > 0.0) || latent > 64.0 || update > 64.0 || draw > 64.0) {
    set_error("bsg_set_persistent_lags: lags must lie in (0, 64]");
    return BSG_ERR_ARG;
  }
  ctx->plag[0] = latent;
  ctx->plag[1] = update;
  ctx->plag[2] = draw;
  ctx->pplan_ok = false;
  return BSG_OK;
}

int bsg_set_overlap(bsg_ctx *ctx, int on)
{
  CTXCHK(ctx);
  ctx->overlap = on != 0;
  return BSG_OK;
}

int bsg_get_overlap(bsg_ctx *ctx, int *on)
{
  CTXCHK(ctx);
  if (!on) {
    set_error("bsg_get_overlap: NULL output");
    return BSG_ERR_ARG;
  }
  *on = use_overlap(ctx) ? 1 : 0;
  return BSG_OK;
}

// ---- autotuning of the Z-free latent kernel ----------------------------
// Whole sweeps at each lane split, overlapped where the split allows it:
// the split also decides whether the latent launch can carry the Wishart
// draw and whether the sweep overlaps, which can matter more than the latent
// kernel's own time. Every candidate starts from the same saved state, and
// the state (coefficients, X'Z and its pieces, Lambda, pair sums, b, flags,
// sweep index) is put back afterwards, so the chain is not changed.
int bsg_autotune(bsg_ctx *ctx, int reps, int *best_g, double *ms_per_g)
{
  CTXCHK(ctx);
  if (ctx->opts.z_store != 2 || reps < 1) {
    set_error("bsg_autotune: tunes the Z-free sweep (z_store 2); reps >= 1");
    return BSG_ERR_ARG;
  }
  bsg::DevProblem &pb = ctx->pb;
  const size_t C = ctx->C, V = ctx->L.totvox, P = ctx->P;
  struct Part {
    void *p;
    size_t bytes;
  };
  const Part parts[] = {{pb.beta, 4 * C * V * P}, {pb.M, 4 * C * V * P},   {pb.lam, 4 * C * P * P},
                        {pb.lam64, 8 * C * P * P}, {pb.tsum, 8 * C * P * P}, {pb.bwm, 4 * C},
                        {pb.flags, 16},            {pb.mpart, pb.mpart ? 4 * C * pb.mp_stride : 0}};
  size_t total = 0;
  for (const Part &q : parts) total += (q.bytes + 255) & ~(size_t)255;
  char *save = nullptr;
  if (hipMalloc((void **)&save, total) != hipSuccess) {
    set_error("bsg_autotune: hipMalloc of the saved state failed");
    return BSG_ERR_NOMEM;
  }
  auto copy = [&](bool out) {
    size_t off = 0;
    for (const Part &q : parts) {
      if (q.bytes &&
          hipMemcpyAsync(out ? save + off : q.p, out ? q.p : save + off, q.bytes, hipMemcpyDeviceToDevice,
                         ctx->stream) != hipSuccess)
        return false;
      off += (q.bytes + 255) & ~(size_t)255;
    }
    return true;
  };
  const uint32_t sw = ctx->sweep;
  const int tp = ctx->tail_pending, keep = pb.lat_g;
  auto restore = [&]() {
    ctx->sweep = sw;
    ctx->tail_pending = tp;
    return copy(false);
  };
  const int cand[4] = {1, 2, 4, 8};
  double best = 1e300;
  int bg = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = copy(true) && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess ? BSG_OK : BSG_ERR_HIP;
  for (int k = 0; k < 4 && rc == BSG_OK; k++) {
    const int G = cand[k];
    double ms = -1.0;
    if (pb.xbx && G != 4) {  // xtz_mode 3 runs at G = 4 only
      if (ms_per_g) ms_per_g[k] = ms;
      continue;
    }
    {
      pb.lat_g = G;
      float f = 0.0f;
      if (!restore() || sweep_run(ctx, 1, nullptr) != BSG_OK || !restore() ||
          hipEventRecord(e0, ctx->stream) != hipSuccess || sweep_run(ctx, reps, nullptr) != BSG_OK ||
          hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&f, e0, e1) != hipSuccess) {
        rc = BSG_ERR_HIP;
        break;
      }
      ms = f / reps;
      if (ms < best) {
        best = ms;
        bg = G;
      }
    }
    if (ms_per_g) ms_per_g[k] = ms;
  }
  if (!restore() || hipStreamSynchronize(ctx->stream) != hipSuccess) rc = BSG_ERR_HIP;
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipFree(save);
  if (rc != BSG_OK) {
    pb.lat_g = keep;
    set_error("bsg_autotune: a timed sweep failed");
    return rc;
  }
  if (pb.lat_g != bg || bg != keep) drop_graph(ctx);
  pb.lat_g = bg;
  if (best_g) *best_g = bg;
  return BSG_OK;
}

// ---- batches of independent analyses ------------------------------------
struct bsg_batch {
  std::vector<bsg_ctx *> m;
  // one stream per covariate count present: the members of one P share
  // every launch, and members of different P are independent, so their
  // launches overlap
  hipStream_t sp[bsg::MAX_P + 1] = {};
  bsg::DevProblem *d_pbs = nullptr;
  bsg::WorkItem *d_items = nullptr;
  bsg::Accum *d_accs = nullptr;
  // item ranges per phase (latent, coef 0, coef 1, wishart, latent of the
  // members drawn by inversion at G = 4, latent with Z stored and X'Z on
  // MFMA of the members in the precision form) and P; save sweeps accumulate
  // in the coefficient phases
  int off[7][bsg::MAX_P + 1] = {}, cnt[7][bsg::MAX_P + 1] = {};
  bool zform = false;  // a member in the precision form (one launch per phase and P then)
  size_t lat_lds[bsg::MAX_P + 1] = {};  // per P: the largest X staging of its members
  // one launch per phase over every covariate count, on s0 (phases latent,
  // colour 0, colour 1, Wishart): chosen for batches whose sweep is too small
  // to keep the chip busy between launches (launch_mode 1, or 0 with few draws)
  int off_all[4] = {}, cnt_all[4] = {};
  size_t lds_all = 0;
  long long draws = 0;  // latent draws per batch sweep
  int launch_mode = 0;  // 0 chosen from the draws, 1 one launch per phase, 2 one per phase and P
  bool unified() const { return !zform && (launch_mode == 1 || (launch_mode == 0 && draws < (1LL << 28))); }
  bool bmf = false;
  std::vector<bsg::Accum> h_accs;
  // graph-captured groups of batch sweeps (bsg_batch_sweep_graph): the
  // members' problems again with their Philox sweep word read from gbase,
  // which the graph's last node advances; s0 forks to and joins the P streams
  bsg::DevProblem *d_pbs_g = nullptr;
  uint32_t *gbase = nullptr;
  uint32_t gstage = 0;
  hipStream_t s0 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join[bsg::MAX_P + 1] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_k = 0;

  ~bsg_batch()
  {
    for (hipStream_t s : sp)
      if (s) hipStreamSynchronize(s);
    if (s0) hipStreamSynchronize(s0);
    if (graph_exec) hipGraphExecDestroy(graph_exec);
    if (graph) hipGraphDestroy(graph);
    if (d_pbs) hipFree(d_pbs);
    if (d_pbs_g) hipFree(d_pbs_g);
    if (gbase) hipFree(gbase);
    if (d_items) hipFree(d_items);
    if (d_accs) hipFree(d_accs);
    if (ev_fork) hipEventDestroy(ev_fork);
    for (hipEvent_t e : ev_join)
      if (e) hipEventDestroy(e);
    for (hipStream_t s : sp)
      if (s) hipStreamDestroy(s);
    if (s0) hipStreamDestroy(s0);
  }
};

int bsg_batch_create(int n, bsg_ctx *const *members, bsg_batch **out)
{
  if (!out || !members || n < 1 || n > bsg::MAX_BATCH) {
    set_error("bsg_batch_create: need 1.." + std::to_string(bsg::MAX_BATCH) + " contexts");
    return BSG_ERR_ARG;
  }
  *out = nullptr;
  std::unique_ptr<bsg_batch> b(new (std::nothrow) bsg_batch);
  if (!b) {
    set_error("bsg_batch_create: out of host memory");
    return BSG_ERR_NOMEM;
  }
  std::vector<bsg::DevProblem> pbs;
  std::vector<bsg::WorkItem> lists[7];
  for (int a = 0; a < n; a++) {
    bsg_ctx *c = members[a];
    if (!c) {
      set_error("bsg_batch_create: NULL context");
      return BSG_ERR_ARG;
    }
    if (c->P > bsg::MAX_P) {
      set_error("bsg_batch_create: batches take members of at most " + std::to_string(bsg::MAX_P) + " covariates");
      return BSG_ERR_ARG;
    }
    // the Z-free form, or the precision form (Z stored, X'Z on MFMA, inversion draws)
    const bool zfree = c->opts.z_store == 2 && c->opts.xtz_mode >= 2;
    const bool zprec = c->opts.z_store == 0 && c->opts.xtz_mode == 1 && c->opts.latent_method == 0;
    if (!(zfree || zprec) || c->pb.wm_update || c->L.n[1] == 0) {
      set_error("bsg_batch_create: members must use a Z-free form (xtz_mode 2 or 3, z_store 2) or the precision "
                "form with Z stored (xtz_mode 1, z_store 0, latent_method 0), have no WM image and two colours");
      return BSG_ERR_ARG;
    }
    if (c->opts.latent_method != members[0]->opts.latent_method || c->sweep != members[0]->sweep) {
      set_error("bsg_batch_create: members must share the latent method and the sweep index");
      return BSG_ERR_ARG;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    bsg::DevProblem pb = c->pb;
    if (zfree && pb.lat_g <= 0) pb.lat_g = bsg::latent_vl_groups(pb);
    pbs.push_back(pb);
    if (zfree) b->lat_lds[pb.P] = std::max(b->lat_lds[pb.P], bsg::latent_vl_lds_bytes(pb));
    b->zform = b->zform || zprec;
    const int lat_list = zprec ? 5 : ((pb.lat_g == 4 && c->opts.latent_method == 0) ? 4 : 0);
    // colour-0 items past its blocks write colour 1's m_j from its pieces
    const int combine = (zfree && bsg::vl_ctiles(pb)) ? bsg::vl_range(pb, 1).left : 0;
    const int tiles = zprec ? bsg::latent_blocks(pb)
                            : (lat_list == 4 ? bsg::latent_vl_blocks_g4(pb) : bsg::latent_vl_tiles(pb, pb.lat_g));
    for (int ch = 0; ch < c->C; ch++) {
      for (int t = 0; t < tiles; t++) lists[lat_list].push_back({a, ch, t});
      if (zprec)
        for (int t = 0; t < bsg::xtz_mfma_blocks(pb); t++) lists[6].push_back({a, ch, t});
      for (int k = 0; k < (c->L.n[0] + 255) / 256 + combine; k++) lists[1].push_back({a, ch, k});
      for (int k = 0; k < (c->L.n[1] + 255) / 256; k++) lists[2].push_back({a, ch, k});
      lists[3].push_back({a, ch, 0});
    }
    b->m.push_back(c);
  }
  b->bmf = members[0]->opts.latent_method == 1;
  std::vector<bsg::WorkItem> all;
  for (int k = 0; k < 7; k++)
    for (int P = 1; P <= bsg::MAX_P; P++) {
      b->off[k][P] = (int)all.size();
      for (const bsg::WorkItem &w : lists[k])
        if (pbs[w.a].P == P) all.push_back(w);
      b->cnt[k][P] = (int)all.size() - b->off[k][P];
    }
  // the same items once more, all covariate counts together, per phase
  for (int ph = 0; ph < 4; ph++) {
    b->off_all[ph] = (int)all.size();
    if (ph == 0) {
      for (const bsg::WorkItem &w : lists[4]) all.push_back(w);
      for (const bsg::WorkItem &w : lists[0]) all.push_back(w);
    } else {
      for (const bsg::WorkItem &w : lists[ph]) all.push_back(w);
    }
    b->cnt_all[ph] = (int)all.size() - b->off_all[ph];
  }
  for (int P = 1; P <= bsg::MAX_P; P++) b->lds_all = std::max(b->lds_all, b->lat_lds[P]);
  for (bsg_ctx *c : b->m) b->draws += (long long)c->C * c->N * c->L.totvox;
  if (hipMalloc((void **)&b->d_pbs, sizeof(bsg::DevProblem) * n) != hipSuccess ||
      hipMalloc((void **)&b->d_items, sizeof(bsg::WorkItem) * all.size()) != hipSuccess ||
      hipMalloc((void **)&b->d_accs, sizeof(bsg::Accum) * n) != hipSuccess) {
    set_error("bsg_batch_create: hipMalloc failed");
    return BSG_ERR_NOMEM;
  }
  for (int P = 1; P <= bsg::MAX_P; P++)
    if (b->cnt[0][P] + b->cnt[4][P] + b->cnt[5][P] > 0) {
      HIPCHK(hipStreamCreateWithFlags(&b->sp[P], hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&b->ev_join[P], hipEventDisableTiming));
    }
  HIPCHK(hipStreamCreateWithFlags(&b->s0, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
  HIPCHK(hipMemcpy(b->d_pbs, pbs.data(), sizeof(bsg::DevProblem) * n, hipMemcpyHostToDevice));
  if (hipMalloc((void **)&b->d_pbs_g, sizeof(bsg::DevProblem) * n) != hipSuccess ||
      hipMalloc((void **)&b->gbase, sizeof(uint32_t)) != hipSuccess) {
    set_error("bsg_batch_create: hipMalloc failed");
    return BSG_ERR_NOMEM;
  }
  for (bsg::DevProblem &q : pbs) q.sweep_base = b->gbase;
  HIPCHK(hipMemcpy(b->d_pbs_g, pbs.data(), sizeof(bsg::DevProblem) * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_items, all.data(), sizeof(bsg::WorkItem) * all.size(), hipMemcpyHostToDevice));
  *out = b.release();
  return BSG_OK;
}

int bsg_batch_destroy(bsg_batch *b)
{
  delete b;
  return BSG_OK;
}

// One sweep of every analysis and chain, one launch per phase and covariate
// count P: latent draws and X'Z (the precision form: latent draws with Z
// stored, then X'Z on MFMA), colour 0, colour 1 (+ pair partials), Wishart,
// in that order on the stream of P. The kernels read their Philox
// sweep word as sw + *sweep_base of the problems in pbs.
static int batch_enqueue(bsg_batch *b, const bsg::DevProblem *pbs, uint32_t sw, const bsg::Accum *accs,
                         const bsg::BatchIndex &bidx)
{
  const bsg::WorkItem *it = b->d_items;
  if (b->unified()) {  // four launches on s0
    for (int ph = 0; ph < 4; ph++)
      HIPCHK(bsg::launch_batch_all(ph, pbs, it + b->off_all[ph], b->cnt_all[ph], b->lds_all, b->bmf, sw, b->s0, accs,
                                   bidx));
    return BSG_OK;
  }
  for (int P = 1; P <= bsg::MAX_P; P++) {
    hipStream_t st = b->sp[P];
    if (!st) continue;
    HIPCHK(bsg::launch_latent_multi(P, pbs, it + b->off[0][P], b->cnt[0][P], b->lat_lds[P], b->bmf, false, sw, st));
    HIPCHK(bsg::launch_latent_multi(P, pbs, it + b->off[4][P], b->cnt[4][P], b->lat_lds[P], false, true, sw, st));
    HIPCHK(bsg::launch_latent_multi_z(P, pbs, it + b->off[5][P], b->cnt[5][P], sw, st));
    HIPCHK(bsg::launch_xtz_multi(P, pbs, it + b->off[6][P], b->cnt[6][P], st));
    HIPCHK(bsg::launch_coef_multi(P, pbs, it + b->off[1][P], b->cnt[1][P], 0, false, sw, st, accs, bidx));
    HIPCHK(bsg::launch_coef_multi(P, pbs, it + b->off[2][P], b->cnt[2][P], 1, true, sw, st, accs, bidx));
    HIPCHK(bsg::launch_wishart_multi(P, pbs, it + b->off[3][P], b->cnt[3][P], sw, st));
  }
  return BSG_OK;
}

// Sweeps of every analysis and chain (batch_enqueue per sweep).
static int batch_run(bsg_batch *b, int nsweeps, int accum_every)
{
  // the members' own queued work first; every member at the same sweep index,
  // since the batch draws every member's Philox words from one index
  for (bsg_ctx *c : b->m) {
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->sweep != b->m[0]->sweep) {
      set_error("bsg_batch_sweep: members are at different sweep indices (a member was swept on its own)");
      return BSG_ERR_ARG;
    }
  }
  bsg::BatchIndex bidx{};
  for (int i = 0; i < nsweeps; i++) {
    const uint32_t sw = b->m[0]->sweep;
    // a save sweep: the colour updates accumulate each member's voxels
    const bool save = accum_every > 0 && (i + 1) % accum_every == 0;
    if (save) {
      for (size_t a = 0; a < b->m.size(); a++) {
        bsg_ctx *c = b->m[a];
        bidx.b[a] = acc_batch_index(c);
        c->acc_count++;
      }
    }
    const bsg::Accum *accs = save ? b->d_accs : nullptr;
    const int rc = batch_enqueue(b, b->d_pbs, sw, accs, bidx);
    if (rc != BSG_OK) return rc;
    for (bsg_ctx *c : b->m) c->sweep++;
  }
  for (hipStream_t st : b->sp)
    if (st) HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipStreamSynchronize(b->s0));
  return BSG_OK;
}

int bsg_batch_set_launch_mode(bsg_batch *b, int mode)
{
  if (!b || mode < 0 || mode > 2) {
    set_error("bsg_batch_set_launch_mode: mode must be 0, 1 or 2");
    return BSG_ERR_ARG;
  }
  for (hipStream_t st : b->sp)
    if (st) HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipStreamSynchronize(b->s0));
  if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
  b->graph_k = 0;
  b->launch_mode = mode;
  return BSG_OK;
}

int bsg_batch_get_launch_mode(bsg_batch *b, int *unified)
{
  if (!b || !unified) {
    set_error("bsg_batch_get_launch_mode: bad argument");
    return BSG_ERR_ARG;
  }
  *unified = b->unified() ? 1 : 0;
  return BSG_OK;
}

// A HIP graph of k batch sweeps: s0 forks to every P stream, each runs its
// launches for sweeps 0..k-1 (words read as i + *gbase), s0 joins them and
// advances gbase by k. Captured once per k.
static int batch_build_graph(bsg_batch *b, int k)
{
  if (b->graph_exec && b->graph_k == k) return BSG_OK;
  if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
  HIPCHK(hipStreamBeginCapture(b->s0, hipStreamCaptureModeThreadLocal));
  int rc = hipEventRecord(b->ev_fork, b->s0) == hipSuccess ? BSG_OK : BSG_ERR_HIP;
  for (int P = 1; P <= bsg::MAX_P && rc == BSG_OK; P++)
    if (b->sp[P] && hipStreamWaitEvent(b->sp[P], b->ev_fork, 0) != hipSuccess) rc = BSG_ERR_HIP;
  const bsg::BatchIndex bidx{};
  for (int i = 0; i < k && rc == BSG_OK; i++) rc = batch_enqueue(b, b->d_pbs_g, (uint32_t)i, nullptr, bidx);
  for (int P = 1; P <= bsg::MAX_P && rc == BSG_OK; P++)
    if (b->sp[P] && (hipEventRecord(b->ev_join[P], b->sp[P]) != hipSuccess ||
                     hipStreamWaitEvent(b->s0, b->ev_join[P], 0) != hipSuccess))
      rc = BSG_ERR_HIP;
  if (rc == BSG_OK && bsg::launch_tick(b->gbase, (uint32_t)k, b->s0) != hipSuccess) rc = BSG_ERR_HIP;
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(b->s0, &g);
  if (rc != BSG_OK || ec != hipSuccess) {
    if (g) hipGraphDestroy(g);
    if (rc == BSG_OK) set_error(std::string("batch graph capture: ") + hipGetErrorString(ec));
    return rc == BSG_OK ? BSG_ERR_HIP : rc;
  }
  b->graph = g;
  HIPCHK(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
  b->graph_k = k;
  return BSG_OK;
}

int bsg_batch_sweep_graph(bsg_batch *b, int nsweeps, int sweeps_per_graph)
{
  if (!b || nsweeps < 0 || sweeps_per_graph < 1 || sweeps_per_graph > 256) {
    set_error("bsg_batch_sweep_graph: need nsweeps >= 0 and 1 <= sweeps_per_graph <= 256");
    return BSG_ERR_ARG;
  }
  const int k = sweeps_per_graph, q = nsweeps / k;
  if (q > 0) {
    for (bsg_ctx *c : b->m) {
      HIPCHK(hipStreamSynchronize(c->stream));
      if (c->sweep != b->m[0]->sweep) {
        set_error("bsg_batch_sweep: members are at different sweep indices (a member was swept on its own)");
        return BSG_ERR_ARG;
      }
    }
    const int rc = batch_build_graph(b, k);
    if (rc != BSG_OK) return rc;
    b->gstage = b->m[0]->sweep;
    HIPCHK(hipMemcpyAsync(b->gbase, &b->gstage, sizeof(uint32_t), hipMemcpyHostToDevice, b->s0));
    for (int j = 0; j < q; j++) HIPCHK(hipGraphLaunch(b->graph_exec, b->s0));
    for (bsg_ctx *c : b->m) c->sweep += (uint32_t)(q * k);
    // the staging word is read by the copy above: keep it stable until then
    HIPCHK(hipStreamSynchronize(b->s0));
  }
  return batch_run(b, nsweeps - q * k, 0);
}

int bsg_batch_sweep(bsg_batch *b, int nsweeps)
{
  if (!b || nsweeps < 0) {
    set_error("bsg_batch_sweep: bad argument");
    return BSG_ERR_ARG;
  }
  return batch_run(b, nsweeps, 0);
}

int bsg_batch_sweep_accumulate(bsg_batch *b, int nsweeps, int save_every)
{
  if (!b || nsweeps < 0 || save_every < 1) {
    set_error("bsg_batch_sweep_accumulate: bad argument");
    return BSG_ERR_ARG;
  }
  b->h_accs.clear();
  for (bsg_ctx *c : b->m) {
    if (!c->acc.sum) {
      set_error("bsg_batch_sweep_accumulate: call bsg_accum_reset on every member first");
      return BSG_ERR_ARG;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    b->h_accs.push_back(c->acc);
  }
  HIPCHK(hipMemcpy(b->d_accs, b->h_accs.data(), sizeof(bsg::Accum) * b->h_accs.size(), hipMemcpyHostToDevice));
  return batch_run(b, nsweeps, save_every);
}
