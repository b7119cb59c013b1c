// bsg_internal.h -- shared definitions of libbsglmm (host and device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace bsg {

// Philox counter tags: counter = (index, sweep, tag | sub << 4, chain). The sweep
// index is passed to every kernel as an argument by the host, which owns it.
constexpr uint32_t TAG_LATENT = 1;  // index = voxel * NQ + subject quad
constexpr uint32_t TAG_COEF = 2;    // index = voxel, sub = group of 4 normals
constexpr uint32_t TAG_GAMMA = 3;   // index = Bartlett row, sub = trial
constexpr uint32_t TAG_WNORM = 4;   // index = row * P + column
constexpr uint32_t TAG_BWM = 5;     // index = 0: normal of the WM coefficient draw
constexpr uint32_t TAG_FALLBACK = 6; // index = voxel * NQ + quad, sub = subject in quad: exact redraw

constexpr int PAIR_BLOCKS = 512;    // fixed partial count per chain: deterministic sums
constexpr int BSTAT_SLOTS = 4096;   // per-chain partial slots of the WM-coefficient statistic
constexpr int MAX_P = 8;

// Host copy of the reference layout (main.cu:189-394) plus the colour-sorted
// order used on the device: sorted index s runs over INDX[0].hostVox, then
// INDX[1].hostVox.
struct Layout {
  int nrow = 0, ncol = 0, ndepth = 0, nsize = 0;
  int totvox = 0, totvoxp = 0;
  int n[2] = {0, 0};
  std::vector<int> host_idx, host_idx_sc;
  std::vector<int> vox[2];
  std::vector<unsigned char> nbrs[2];
  std::vector<int> rank_of_sorted;  // sorted s -> hostIdx (mask rank)
  std::vector<int> sc_of_sorted;    // sorted s -> hostIdxSC (dilated-mask rank)
  std::vector<int> nbr6;            // [V][6] sorted neighbours, -1 outside mask
};

int build_layout(const unsigned char *mask, int nrow, int ncol, int ndepth, Layout &out);

// Everything a kernel needs, passed by value.
struct DevProblem {
  int V, N, NS, NQ, NW, P, nchains, n0;
  uint32_t key0, key1;
  int df;
  float trunc_eps, pivot_min;
  const int *nbr6;         // [V][6]
  const uint32_t *ybits;   // [V][NW], bit i of row s = (y_is == 1)
  const float *X;          // [P][NS], zero past N
  const float *xtx;        // [P][P]
  const float *xsum;       // [P]  column sums of X
  const float *wm;         // [V]
  float *Z;                // [chain][V][NS]
  float *beta;             // [chain][V][P]
  float *M;                // [chain][V][P]  X'(Z - b WM)
  float *lam;              // [chain][P][P]
  double *lam64;           // [chain][P][P]
  double *tsum;            // [chain][P][P]
  double *partials;        // [chain][max_parts][P(P+1)/2]
  int max_parts;           // partial slots per chain
  float *bwm;              // [chain]  WM coefficient b
  uint32_t *flags;         // [0] FP64 refactors, [1] non-SPD voxels
  int vb, ve;              // voxel range [vb, ve) of a latent / X'Z launch
  // Philox sweep word = *sweep_base + the kernel's sweep argument. Direct
  // launches point at a word holding 0 and pass the absolute index; a
  // captured graph of K sweeps passes 0..K-1 and ends with k_tick(+K).
  const uint32_t *sweep_base;
  int latent_method;       // Z-free form: 0 inversion for every draw, 1 Box-Muller candidate + exact fallback
  int lat_g;               // Z-free form: lanes per voxel (1, 2, 4, 8); 0 = chosen from N
  // WM coefficient b (covarGPU.cu:588-724): drawn each sweep when wm_update
  int wm_update;
  double wm2;              // N sum_j WM_j^2
  double *bstat;           // [chain][BSTAT_SLOTS] block partials of sum_j WM_j (sum_i Z_ij - X'1 b_j)
  // tail split of the Z-free latent launch at G = 4 (k_latent_vl_mfm): the
  // launch's T mod ncu leftover 64-voxel tiles run as tail_k pieces each
  int ncu;                 // compute units of the device
  int tail_tiles, tail_k;  // leftover tiles of a full-range launch; pieces per tile (<= 1: no split)
  float *mpart;            // [chain][tail_tiles][tail_k][64][P] partial X'Z of the pieces
  int zfree;               // Z not kept, X'Z fused into the latent kernel (z_store 2)
};

// Posterior accumulators (device, per chain): see k_accum.
struct Accum {
  int ntypes = 0, nbatches = 0;
  double *sum = nullptr;   // [chain][V][P]
  double *sq = nullptr;    // [chain][V][P]
  double *prb = nullptr;   // [chain][V][ntypes]
  double *bwm = nullptr;   // [chain][V]
  float *batch = nullptr;  // [chain][nbatches][V][P + ntypes]
};

// One block of a batched launch: analysis, chain and tile (or block) index.
struct WorkItem {
  int a, chain, blk;
};
constexpr int MAX_BATCH = 32;  // analyses per batch
struct BatchIndex {
  int b[MAX_BATCH];  // posterior-accumulator batch of each analysis
};

// kernel launchers (bsg_kernels.hip)
// Returns, in *bstat_parts, how many WM-statistic partials the launch wrote
// (0 when the kernel used does not form them; see launch_bwm).
// wish = 1: the launch also draws the CAR precision of sweep wsweep (block 0;
// only where latent_carries_wishart)
hipError_t launch_latent(const DevProblem &pb, bool fused_xtz, int zstore, uint32_t sweep, hipStream_t st,
                         int v0, int v1, int *bstat_parts = nullptr, int wish = 0, uint32_t wsweep = 0);
bool latent_carries_wishart(const DevProblem &pb);
// WM coefficient draw. parts > 0: the latent kernel already wrote that many
// partials; parts == 0: they are formed here from the stored Z first.
hipError_t launch_bwm(const DevProblem &pb, int parts, uint32_t sweep, hipStream_t st);
hipError_t launch_xtz(const DevProblem &pb, int mode, hipStream_t st, int v0, int v1);
// tail = 1: the last latent launch of this problem was split (latent_split_active):
// the update adds its own voxels' pieces, and colour 0 also writes colour 1's
// m_j from their pieces in extra blocks, so colour 1 then runs with tail = 0
// ac != null: a save sweep, each updated voxel is accumulated into ac (batch)
hipError_t launch_coef(const DevProblem &pb, int colour, bool with_prec, uint32_t sweep, hipStream_t st, int tail = 0,
                       const Accum *ac = nullptr, int batch = 0);
bool latent_split_active(const DevProblem &pb, int *nmain, int *K);
hipError_t launch_pairs(const DevProblem &pb, hipStream_t st);
hipError_t launch_wishart(const DevProblem &pb, bool side, uint32_t sweep, hipStream_t st);
hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st);
hipError_t launch_probe(int kind, const void *src, void *dst, size_t bytes, hipStream_t st);
hipError_t launch_accum(const DevProblem &pb, const Accum &ac, int batch, hipStream_t st);
hipError_t launch_tick(uint32_t *base, uint32_t k, hipStream_t st);
// Z-free latent form: lanes per voxel chosen from N, its tile count and LDS bytes
int latent_vl_groups(const DevProblem &pb);
int latent_vl_tiles(const DevProblem &pb, int G);
int latent_vl_blocks_g4(const DevProblem &pb);  // blocks of a full-range G = 4 inversion launch (tail split)
size_t latent_vl_lds_bytes(const DevProblem &pb);
// batched analyses: one launch per phase and covariate count P over the work
// items; latent items of members at G = 4 (inversion) take their own launch (g4)
hipError_t launch_latent_multi(int P, const DevProblem *pbs, const WorkItem *items, int n, size_t lds, bool bmf,
                               bool g4, uint32_t sweep, hipStream_t st);
// accs != null: a save sweep, each member's updated voxels are accumulated (batch bidx)
hipError_t launch_coef_multi(int P, const DevProblem *pbs, const WorkItem *items, int n, int colour, bool prec,
                             uint32_t sweep, hipStream_t st, const Accum *accs, const BatchIndex &bidx);
hipError_t launch_wishart_multi(int P, const DevProblem *pbs, const WorkItem *items, int n, uint32_t sweep,
                                hipStream_t st);
// plane means: part[chain][blocks][P] FP64 block partials
hipError_t launch_cmeans(const DevProblem &pb, double *part, int blocks, hipStream_t st);
// per-subject log-likelihood: part [chain][tiles][N][T], out [chain][N][T]
int loglik_tiles(const DevProblem &pb);
hipError_t launch_loglik(const DevProblem &pb, int T, double *part, double *out, hipStream_t st);

void set_error(const std::string &msg);

}  // namespace bsg
