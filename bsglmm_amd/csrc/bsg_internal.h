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
constexpr uint32_t TAG_PROBE = 7;    // the sampler probe's own stream (no chain state)

constexpr int PAIR_BLOCKS = 512;    // fixed partial count per chain: deterministic sums
constexpr int BSTAT_SLOTS = 4096;   // per-chain partial slots of the WM-coefficient statistic
constexpr int MAX_P = 8;      // covariates of the register-resident kernels (and of batches)
constexpr int PW_MAX = 20;    // covariates of the wide kernels (the reference's limit, covarGPU.cu:65)

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
  const float4 *Xq;        // [NQ][P] quads of X, zero past N (wide problems, P > MAX_P)
  const uint16_t *xbf;     // [2P][NSB] BF16 bits of X_hi (rows 0..P-1) and X_lo (P..2P-1), zero past N (P <= MAX_P)
  int NSB;                 // NS rounded up to a multiple of 8
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
  // colour tiling of the Z-free latent launch at G = 4 (k_latent_vl_mfm, one
  // chain): each colour's voxels in 64-voxel tiles from its first voxel; a
  // colour's T mod ncu leftover tiles run as tl_k pieces each when tl_k > 1
  int ncu;                 // compute units of the device
  int ctile;               // colour tiling allowed (one chain)
  int tl_left[2], tl_k[2]; // per colour: leftover tiles, pieces per leftover tile (<= 1: no split)
  float *mpart;            // [chain][colour 0 pieces | colour 1 pieces][64][P] partial X'Z of the pieces
  int mp_stride;           // floats of mpart per chain
  int zfree;               // Z not kept, X'Z fused into the latent kernel (z_store 2)
  int xbx;                 // the fused X'Z as BF16 hi + lo products on the matrix cores (xtz_mode 3, G = 4)
  // Lambda of the overlapped sweep: (1 << 32 | float bits) per entry, written
  // by the block that draws it, read by the colour-0 blocks of the same
  // launch once their entries show the tag (no fence needed: each entry
  // carries its tag); cleared by the launch before
  unsigned long long *lamtag;  // [chain][P][P]
  double *wbart;               // [chain][P][P] Bartlett factor of the next overlapped Lambda draw
  // bounded wait for the tagged Lambda: polls before giving up (0: 2^20), and
  // the tag the draw writes (0: 1, the one the readers wait for). Set from the
  // host only to test the bound (bsg_test_overlap_wait).
  int lam_spin, lam_tag;
  // diagnosis of the persistent sweep's waits (null outside it), nb blocks:
  // [4 + b] the item block b runs (k << 32 | queue entry), [4 + nb + b] the
  // item it last recorded, then from [4 + 2 nb] four words per wait past its
  // bound (one per item of a block; flags[3] counts them): the word's
  // address, the tag waited for, the word read last, the item.
  unsigned long long *pdiag;
};

// The persistent sweep (k_persist, one chain): every item of K sweeps in one
// launch, taken from a work queue in an order that puts each item after
// everything it reads. Values passed between the blocks of the launch travel
// as tagged words: the tag (high 32 bits) names the relative sweep that wrote
// the payload (low 32 bits), so a reader waits for its own tag and needs no
// fence; FP64 values take two words (high half, low half).
// Item types of the queue pattern (entry = dk << 28 | type << 24 | index):
enum PItem : uint32_t { PI_L0 = 0, PI_L1 = 1, PI_U0 = 2, PI_U1 = 3, PI_W = 4, PI_BT = 5 };
constexpr int PERSIST_MAX_K = 65534;  // sweeps per launch (16-bit relative sweep in the tag)
constexpr int PERSIST_MAX_EPOCH = 32767;  // launches before the tagged words are cleared
struct PTags {
  unsigned long long *bt;   // [2][V][P] coefficients; buffer k & 1 written by the updates of relative sweep k
  unsigned long long *mt;   // [V][P] X'Z of whole tiles
  unsigned long long *mpt;  // partial X'Z of the pieces, laid out as mpart (one chain)
  unsigned long long *pt;   // [colour-1 blocks][2 NE] pair partials (FP64)
  unsigned long long *lt;   // [2][P][P] Lambda drawn at relative sweep k, buffer k & 1
  unsigned long long *wb;   // [2][P][P][2] Bartlett factor (FP64) of relative sweep k, buffer k & 1
  unsigned long long *u0d;  // [colour-0 blocks] the colour-0 update of the block is done
  uint32_t *head;           // next queue entry (zero at launch)
  const uint32_t *pattern;  // one window of the queue: plen entries
  int plen, windows, K;     // entries per window, windows, sweeps of the launch
  uint32_t epoch;           // 1 .. PERSIST_MAX_EPOCH: tag = epoch << 16 | (k + 1)
  uint32_t sweep0;          // Philox sweep word of relative sweep 0, before *sweep_base
  int tail;                 // bit c: colour c has leftover tiles in pieces
};

// Lanes per voxel of the Z-free latent kernels. Inversion draws take G = 4 at
// every N: there the means run on the matrix cores, the latent launch carries
// the Wishart draw and the sweep overlaps (at small N this doubles the sweep
// rate, the kernel's own time aside). The accept-or-fallback draws take
// enough lanes per voxel to give the chip many waves, few enough that every
// lane keeps several quads.
__host__ __device__ inline int vl_groups(const DevProblem &pb)
{
  if (pb.lat_g > 0) return pb.lat_g;
  if (pb.latent_method == 0) return 4;
  if (pb.NQ <= 4) return 1;
  if (pb.NQ <= 16) return 2;
  if (pb.NQ <= 128) return 4;
  return 8;
}

// Colour tiling applies: one chain, Z-free inversion draws at G = 4, b fixed
__host__ __device__ inline bool vl_ctiles(const DevProblem &pb)
{
  return pb.ctile && pb.zfree && pb.latent_method == 0 && !pb.wm_update && vl_groups(pb) == 4;
}

// Tiles of colour c (sorted voxels [base, end)): nmain whole tiles, then
// `left` leftover tiles of K pieces each, whose partial X'Z sums sit at
// mpart + chain * mp_stride + moff (piece k of leftover tile t at (t K + k) 64 P)
struct VlRange {
  int base, end, nmain, left, K, moff;
};
__host__ __device__ inline VlRange vl_range(const DevProblem &pb, int c)
{
  VlRange r;
  r.base = c ? pb.n0 : 0;
  r.end = c ? pb.V : pb.n0;
  const int tiles = (r.end - r.base + 63) / 64;
  r.left = pb.tl_k[c] > 1 ? pb.tl_left[c] : 0;
  r.K = r.left ? pb.tl_k[c] : 1;
  r.nmain = tiles - r.left;
  r.moff = (c && pb.tl_k[0] > 1) ? pb.tl_left[0] * pb.tl_k[0] * 64 * pb.P : 0;
  return r;
}
// blocks of a colour-tiled launch over colour set cs (bit c: colour c)
__host__ __device__ inline int vl_blocks(const DevProblem &pb, int cs)
{
  int n = 0;
  for (int c = 0; c < 2; c++)
    if ((cs >> c) & 1) {
      const VlRange r = vl_range(pb, c);
      n += r.nmain + r.left * r.K;
    }
  return n;
}

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
// k_latent (Z stored): lanes per voxel 2^lshift, the fewest covering the
// subject quads up to 64, and reps quads per lane; each wave walks
// LATENT_VITER voxel groups
constexpr int LATENT_VITER = 8;
__host__ __device__ inline void lane_grouping(const DevProblem &pb, int &lshift, int &reps)
{
  int L = 1;
  lshift = 0;
  while (L < pb.NQ && L < 64) {
    L <<= 1;
    lshift++;
  }
  reps = (pb.NQ + L - 1) / L;
}
// blocks of a whole-range k_latent launch (one chain) and of k_xtz_mfma
inline int latent_blocks(const DevProblem &pb)
{
  int lshift, reps;
  lane_grouping(pb, lshift, reps);
  const int vpw = 64 >> lshift;
  const long long groups = ((long long)pb.V + vpw - 1) / vpw;
  const long long waves = (groups + LATENT_VITER - 1) / LATENT_VITER;
  return (int)((waves + 3) / 4);
}
inline int xtz_mfma_blocks(const DevProblem &pb)
{
  const long long tiles = ((long long)pb.V + 15) / 16;
  return (int)((tiles + 3) / 4);
}
constexpr int MAX_BATCH = 32;  // analyses per batch
struct BatchIndex {
  int b[MAX_BATCH];  // posterior-accumulator batch of each analysis
};

// kernel launchers (bsg_kernels.hip)
// Returns, in *bstat_parts, how many WM-statistic partials the launch wrote
// (0 when the kernel used does not form them; see launch_bwm).
// wish = 1: the launch also draws the CAR precision of sweep wsweep (block 0;
// only where latent_carries_wishart)
// A colour-tiled launch (vl_ctiles) over the whole range, or over one colour's
// voxels [0, n0) / [n0, V), writes its leftover tiles' X'Z as pieces:
// *pieces (optional) gets the colour set whose m_j then sit in mpart.
hipError_t launch_latent(const DevProblem &pb, bool fused_xtz, int zstore, uint32_t sweep, hipStream_t st,
                         int v0, int v1, int *bstat_parts = nullptr, int wish = 0, uint32_t wsweep = 0,
                         int *pieces = nullptr);
bool latent_carries_wishart(const DevProblem &pb);
// Overlapped sweep (every chain; one chain runs colour tiles and pieces;
// Lambda drawn inside the launch):
// phase 0: [Wishart of sweep wsweep when wish] colour 1's latent draws of
//          `sweep`, then colour 0's update of `sweep` (reading the new Lambda
//          by its tags when wish); `lat` is ignored;
// phase 1: colour 0's latent draws of sweep + 1 (when lat), colour 1's
//          update of `sweep` (pair partials), then the Bartlett factor of
//          `sweep`'s Lambda for the next phase 0.
// tail: colour set whose leftover pieces the update adds (its own colour).
bool overlap_supported(const DevProblem &pb);
hipError_t launch_overlap(const DevProblem &pb, int phase, uint32_t sweep, int lat, int wish, uint32_t wsweep,
                          int tail, hipStream_t st, const Accum *ac = nullptr, int batch = 0);
// The persistent sweep: the overlapped form's conditions with one chain (colour
// tiles) and P <= 6. launch_persist clears the queue head, then runs nblocks
// persistent blocks over tg.windows windows of the pattern.
bool persist_supported(const DevProblem &pb);
hipError_t launch_persist(const DevProblem &pb, const PTags &tg, int nblocks, hipStream_t st);
int persist_max_blocks(const DevProblem &pb);  // resident blocks of the kernel on this device
// zero n tagged words (a fill kernel: no memset node if captured)
hipError_t launch_zero64(unsigned long long *p, size_t n, hipStream_t st);
// the BF16 hi/lo split of X for the MFMA X'Z (pb.xbf), from pb.X, once at creation
hipError_t launch_xbf(const DevProblem &pb, uint16_t *xbf, hipStream_t st);
// WM coefficient draw. parts > 0: the latent kernel already wrote that many
// partials; parts == 0: they are formed here from the stored Z first.
hipError_t launch_bwm(const DevProblem &pb, int parts, uint32_t sweep, hipStream_t st);
hipError_t launch_xtz(const DevProblem &pb, int mode, hipStream_t st, int v0, int v1);
// tail: colour set (bit c) whose leftover-tile m_j are still pieces in mpart.
// The update adds its own colour's pieces; colour 0 also writes colour 1's
// m_j from theirs in extra blocks, so colour 1 then finds them whole.
// ac != null: a save sweep, each updated voxel is accumulated into ac (batch)
hipError_t launch_coef(const DevProblem &pb, int colour, bool with_prec, uint32_t sweep, hipStream_t st, int tail = 0,
                       const Accum *ac = nullptr, int batch = 0);
hipError_t launch_pairs(const DevProblem &pb, hipStream_t st);
hipError_t launch_wishart(const DevProblem &pb, bool side, uint32_t sweep, hipStream_t st);
hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st);
hipError_t launch_probe(int kind, const void *src, void *dst, size_t bytes, hipStream_t st);
constexpr int VALU_CHAINS_PER_ROUND = 8;  // vector instructions per wave and round of k_valu
hipError_t launch_valu(int kind, float *out, int blocks, int iters, hipStream_t st);
// The latent sampler alone (Philox4x32-10 and the inversion draw of a quad,
// no loads or stores in the loop): blocks of 256 threads, `rounds` quads per
// thread; out holds at least `blocks` floats.
hipError_t launch_sampler_probe(float *out, int blocks, int rounds, hipStream_t st);
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
// precision-form members: Z stored (k_latent body) and X'Z on MFMA (k_xtz_mfma body)
hipError_t launch_latent_multi_z(int P, const DevProblem *pbs, const WorkItem *items, int n, uint32_t sweep,
                                 hipStream_t st);
hipError_t launch_xtz_multi(int P, const DevProblem *pbs, const WorkItem *items, int n, hipStream_t st);
hipError_t launch_coef_multi(int P, const DevProblem *pbs, const WorkItem *items, int n, int colour, bool prec,
                             uint32_t sweep, hipStream_t st, const Accum *accs, const BatchIndex &bidx);
hipError_t launch_wishart_multi(int P, const DevProblem *pbs, const WorkItem *items, int n, uint32_t sweep,
                                hipStream_t st);
// batches in one launch per phase over every covariate count (phase 0 latent,
// 1 colour 0, 2 colour 1 with the pair partials, 3 Wishart); lds: the
// largest latent staging of the members
hipError_t launch_batch_all(int phase, const DevProblem *pbs, const WorkItem *items, int n, size_t lds, bool bmf,
                            uint32_t sweep, hipStream_t st, const Accum *accs, const BatchIndex &bidx);
// plane means: part[chain][blocks][P] FP64 block partials
hipError_t launch_cmeans(const DevProblem &pb, double *part, int blocks, hipStream_t st);
// out[c] = sum of part[b][c] over b < blocks in block order (one chain)
hipError_t launch_cmeans_fold(const double *part, int blocks, int P, double *out, hipStream_t st);

// Allocation-free scratch: a fixed device array of SCRATCH_DOUBLES doubles
// (64 KiB) per device. scratch_init takes its address on the current device
// (once per device; bsg_create calls it); scratch_ptr returns that address,
// or null before scratch_init ran on the current device.
constexpr size_t SCRATCH_DOUBLES = 8192;
constexpr int MAX_DEVICES = 64;
hipError_t scratch_init();
double *scratch_ptr();
// per-subject log-likelihood: part [chain][tiles][N][T], out [chain][N][T]
int loglik_tiles(const DevProblem &pb);
hipError_t launch_loglik(const DevProblem &pb, int T, double *part, double *out, hipStream_t st);

void set_error(const std::string &msg);

}  // namespace bsg
