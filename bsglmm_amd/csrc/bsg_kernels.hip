// bsg_kernels.hip -- the Gibbs sweep kernels for gfx950 (MI355X).
//
// One sweep (reference order mcmc.cpp:311-394, probit, no WM image):
//   k_latent   Z_ij ~ TN(x_i'b_j + b WM_j, 1) on (eps,500) or (-500,-eps)
//              (replaces Z_GPU1, covarGPU.cu:1368-1428)
//   k_xtz      m_j = X'(Z_j - b WM_j)                  (covarGPU.cu:327-352)
//   k_coef     checkerboard b_j update, one colour      (covarGPU.cu:354-392)
//   k_pairs    FP64 sums of neighbour-pair outer products (covarGPU.cu:953-1013)
//   k_wishart  Lambda ~ Wishart(V+19, (T+I)^-1) on the device
//              (covarGPU.cu:1086-1108, cholesky.cpp:245-286)
//
// Layout in HBM (sorted order s: colour-0 list, then colour-1 list):
//   Z[chain][s][NS] voxel-major, NS = N rounded up to 4 (float4 rows)
//   ybits[s][NW]    bit i = (y_is == 1)
//   beta, M [chain][s][P];  nbr6[s][6] sorted neighbour indices.
//
// Random variates come from stateless Philox4x32-10 with counter
// (index, sweep, tag | sub << 4, chain); no per-thread generator state.
#include "bsg_internal.h"

// scope of the tagged words' loads and stores
#define TAG_SCOPE __HIP_MEMORY_SCOPE_AGENT

#include <algorithm>
#include <mutex>

namespace bsg {

// ---------------------------------------------------------------------------
// Philox4x32-10 and uniforms. The 32x32->64 products are written as 64-bit
// multiplies so each becomes one v_mad_u64_u32 (high and low halves at once).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);  // v_bitop3_b32, truth table of a^b^c
}

__device__ __forceinline__ void philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                         uint32_t k0, uint32_t k1, uint32_t r[4])
{
#pragma unroll
  for (int i = 0; i < 10; i++) {
    const uint64_t p0 = (uint64_t)c0 * 0xD2511F53u;
    const uint64_t p1 = (uint64_t)c2 * 0xCD9E8D57u;
    c0 = xor3((uint32_t)(p1 >> 32), c1, k0);
    c1 = (uint32_t)p1;
    c2 = xor3((uint32_t)(p0 >> 32), c3, k1);
    c3 = (uint32_t)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  r[0] = c0;
  r[1] = c1;
  r[2] = c2;
  r[3] = c3;
}

// odd multiples of 2^-24: exact in float and double, never 0 or 1.
// (1 + k 2^-23) - (1 - 2^-24) = (2k + 1) 2^-24 exactly (Sterbenz), k = x >> 9.
// v_alignbit_b32(0x7F, x, 9) = (0x7F << 23) | (x >> 9) = 0x3F800000 | (x >> 9)
// in one instruction.
__device__ __forceinline__ float u01f(uint32_t x)
{
  return __uint_as_float(__builtin_amdgcn_alignbit(0x7Fu, x, 9)) - 0.99999994039535522f;
}
__device__ __forceinline__ double u01d(uint32_t x)
{
  return (double)(((x >> 9) << 1) | 1u) * 5.9604644775390625e-8;
}

// ---------------------------------------------------------------------------
// Deterministic block reductions in FP64 (fixed order, so sums repeat bit for bit).
// ---------------------------------------------------------------------------
// Block sum of NE doubles in fixed order; thread e < NE of the block gets entry e.
// (count < NE: only entries e < count are written)
template <int NE>
__device__ __forceinline__ void block_sum_store(double (&acc)[NE], double *dst, int count = NE)
{
  __shared__ double red[4][NE];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int e = 0; e < NE; e++) {
    double v = acc[e];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wv][e] = v;
  }
  __syncthreads();
  if (threadIdx.x < NE && (int)threadIdx.x < count) {
    const int e = threadIdx.x;
    dst[e] = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
  }
}

// block_sum_store's sums, each written as two tagged words (high, low half)
template <int NE>
__device__ __forceinline__ void block_sum_store_tg(double (&acc)[NE], unsigned long long *dst, uint32_t tag)
{
  __shared__ double red[4][NE];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int e = 0; e < NE; e++) {
    double v = acc[e];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wv][e] = v;
  }
  __syncthreads();
  if (threadIdx.x < NE) {
    const int e = threadIdx.x;
    const double v = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    __hip_atomic_store(dst + 2 * e, ((unsigned long long)tag << 32) | (b >> 32), __ATOMIC_RELAXED,
                       TAG_SCOPE);
    __hip_atomic_store(dst + 2 * e + 1, ((unsigned long long)tag << 32) | (b & 0xFFFFFFFFull), __ATOMIC_RELAXED,
                       TAG_SCOPE);
  }
}

// Sum of the per-block partials in a fixed order: thread t adds blocks
// t, t + 256, ... of an entry, then a block reduction (block_sum_store), for
// all entries at once up to P = 6 and in chunks of 6 above (each entry's sum
// is formed alone, so the chunking changes no result; it keeps the
// accumulators within a latent kernel's register budget at P = 7, 8). T (shared memory) holds the sums for
// every thread on return.
template <int NE>
__device__ void sum_partials(int nparts, const double *part, double *T)
{
  constexpr int CH = NE <= 21 ? NE : 6;  // one chunk up to P = 6
#pragma unroll 1
  for (int e0 = 0; e0 < NE; e0 += CH) {
    double acc[CH];
#pragma unroll
    for (int e = 0; e < CH; e++) acc[e] = 0.0;
    for (int b = threadIdx.x; b < nparts; b += 256) {
#pragma unroll
      for (int e = 0; e < CH; e++)
        if (e0 + e < NE) acc[e] += part[(size_t)b * NE + e0 + e];
    }
    block_sum_store<CH>(acc, T + e0, NE - e0);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Normal tail function and its inverse in FP32, branch free. Coefficients are
// this project's own fits (bsglmm_amd/csrc/fit_normal_tail.py, kernel form):
//   Q(t)    = r 2^(G(r) - t^2 log2(e)/2), r = 1/(1 + t/(2 sqrt 2)),
//             G = log2(e) g with g of degree 7 (fitted for t <= 11.3; past it
//             Q falls below 1e-29 and underflows to 0 by t = 14)
//   Qinv(p) = n(s)/d(s), s = sqrt(-log2 p), n of degree 4, d of degree 3,
//             p in (1e-38, 1/2]
// FP32 errors: Q 6e-6 relative (t <= 8, from rounding the t^2 term), Qinv
// 6e-7 relative (floor 1).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float q_tail(float t)
{
  const float r = __builtin_amdgcn_rcpf(fmaf(0.35355339059327373f, t, 1.0f));
  float g = -0.2613292634487152f;
  g = fmaf(g, r, 1.0420771837234497f);
  g = fmaf(g, r, -1.3670768737792969f);
  g = fmaf(g, r, 0.42922380566596985f);
  g = fmaf(g, r, -0.014621351845562458f);
  g = fmaf(g, r, 0.5545806884765625f);
  g = fmaf(g, r, 1.4430041313171387f);
  g = fmaf(g, r, -2.825859308242798f);
  return r * __builtin_amdgcn_exp2f(fmaf(-0.72134752044448170f * t, t, g));
}

__device__ __forceinline__ float q_inv(float p)
{
  const float s = __builtin_amdgcn_sqrtf(-__builtin_amdgcn_logf(p));
  float n = 0.686313807964325f;
  n = fmaf(n, s, 5.580847263336182f);
  n = fmaf(n, s, 4.647900104522705f);
  n = fmaf(n, s, -7.771159648895264f);
  n = fmaf(n, s, -3.143899917602539f);
  float d = 0.5826886892318726f;
  d = fmaf(d, s, 4.7574782371521f);
  d = fmaf(d, s, 5.753643989562988f);
  d = fmaf(d, s, 1.0f);
  return n * __builtin_amdgcn_rcpf(d);
}

// ---------------------------------------------------------------------------
// Truncated-normal latent draw by inversion (one uniform, no rejection loop).
// W ~ N(0,1) truncated below at a; Z = mu + s W with s = +1 (y=1), -1 (y=0)
// and a = eps - s mu (covarGPU.cu:1414-1422). W = Qinv(u Q(a)); for a < 0,
// Q(a) = 1 - Q(|a|) and, when u Q(a) > 1/2, W = -Qinv(1 - u Q(a)) with
// 1 - u Q(a) = (1-u) + u Q(|a|) formed without cancellation. The upper limit
// 500 (covarGPU.cu:1415,1418) has Q(500 - s mu) = 0 in FP32 for |mu| < 480 and
// is not evaluated. a >= 8 takes a log-space Newton path (rare).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float logq_asym(float x)  // log Q(x) for x >= 8
{
  float r = 1.0f / (x * x);
  float s = 1.0f + r * (-1.0f + r * (3.0f + r * (-15.0f + r * (105.0f + r * -945.0f))));
  return -0.5f * x * x - __logf(x) - 0.918938533204673f + __logf(s);
}

__device__ __noinline__ float tail_w(float a, float u)
{
  float lu = logf(u);
  float lp = lu + logq_asym(a);
  float x = a - lu / a;
#pragma unroll 1
  for (int it = 0; it < 5; it++) {
    float g = logq_asym(x) - lp;
    x += g / (x + 1.0f / x);
  }
  return x;
}

// Four draws of one Philox block: subject l uses uniform l and bit l of yq.
__device__ __forceinline__ void latent_quad(const float (&mu)[4], uint32_t yq, const uint32_t (&rr)[4],
                                            float eps, float (&z)[4])
{
  float a[4], uu[4];
  bool tail = false;
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const float u = u01f(rr[l]);
    const float s = ((yq >> l) & 1u) ? 1.0f : -1.0f;
    const float al = fmaf(-s, mu[l], eps);
    const float qt = q_tail(fabsf(al));  // Q(|a|) <= 1/2
    const float uq = u * qt;
    const float pn = u - uq;             // u Q(a) when a < 0
    const bool neg = al < 0.0f;
    const bool flip = neg && pn > 0.5f;
    const float arg = neg ? (flip ? (1.0f - u) + uq : pn) : uq;
    float w = q_inv(arg);
    w = flip ? -w : w;
    w = (w < al) ? al : w;
    z[l] = fmaf(s, w, mu[l]);
    a[l] = al;
    uu[l] = u;
    tail |= al >= 8.0f;
  }
  if (tail) {
#pragma unroll 1
    for (int l = 0; l < 4; l++)
      if (a[l] >= 8.0f) {
        const float s = ((yq >> l) & 1u) ? 1.0f : -1.0f;
        z[l] = fmaf(s, fmaxf(tail_w(a[l], uu[l]), a[l]), mu[l]);
      }
  }
}

// ---------------------------------------------------------------------------
// Cross-lane sum of P per-lane values over a full wave64 (reduce-scatter):
// v_permlane32_swap pairs values (c, c + PP/2) across the wave halves,
// v_permlane16_swap pairs them across 16-lane rows, DPP finishes inside a row
// (or half row for PP = 8). The total of value c ends in lane red_lane<P>(c).
// PP = P padded to 4 or 8 with zeros.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void swap32(float &x, float &y)
{
  auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
  x = __uint_as_float(r[0]);
  y = __uint_as_float(r[1]);
}
__device__ __forceinline__ void swap16(float &x, float &y)
{
  auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
  x = __uint_as_float(r[0]);
  y = __uint_as_float(r[1]);
}
template <int CTRL>
__device__ __forceinline__ float dpp_add(float v)
{
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

template <int P>
constexpr int red_pad() { return P <= 4 ? 4 : 8; }

template <int P>
__device__ __forceinline__ int red_lane(int c)
{
  if constexpr (red_pad<P>() == 4) return 16 * c;
  // rows hold {a0,a2},{a1,a3},{a4,a6},{a5,a7} as (low half, high half)
  const int row = (c & 1) + 2 * (c >> 2);
  const int half = (c >> 1) & 1;
  return 16 * row + 8 * half;
}

// inverse of red_lane: the covariate whose total lane `lane` writes, or -1
template <int P>
__device__ __forceinline__ int red_cov(int lane)
{
  if ((lane & (red_pad<P>() == 4 ? 15 : 7)) != 0) return -1;
  int c;
  if constexpr (red_pad<P>() == 4) {
    c = lane >> 4;
  } else {
    const int row = lane >> 4, half = (lane >> 3) & 1;
    c = (row & 1) + 2 * half + 4 * (row >> 1);
  }
  return c < P ? c : -1;
}

template <int P>
__device__ __forceinline__ float wave_reduce_scatter(const float (&acc)[P], int lane)
{
  constexpr int PP = red_pad<P>();
  float a[PP];
#pragma unroll
  for (int c = 0; c < PP; c++) a[c] = c < P ? acc[c] : 0.0f;
  float b[PP / 2];
#pragma unroll
  for (int k = 0; k < PP / 2; k++) {
    float x = a[k], y = a[k + PP / 2];
    swap32(x, y);
    b[k] = x + y;  // lanes 0-31: a_k, lanes 32-63: a_{k+PP/2}
  }
  if constexpr (PP == 4) {
    float x = b[0], y = b[1];
    swap16(x, y);
    float c = x + y;  // row r holds a_r
    c = dpp_add<0x128>(c);  // row_ror:8
    c = dpp_add<0x124>(c);  // row_ror:4
    c = dpp_add<0x122>(c);  // row_ror:2
    c = dpp_add<0x121>(c);  // row_ror:1
    return c;
  } else {
    float x0 = b[0], y0 = b[1], x1 = b[2], y1 = b[3];
    swap16(x0, y0);
    swap16(x1, y1);
    float c0 = dpp_add<0x128>(x0 + y0);  // rows: a0, a1, a4, a5
    float c1 = dpp_add<0x128>(x1 + y1);  // rows: a2, a3, a6, a7
    float e = (lane & 8) ? c1 : c0;
    e = dpp_add<0x141>(e);  // row_half_mirror
    e = dpp_add<0xB1>(e);   // quad_perm [1,0,3,2]
    e = dpp_add<0x4E>(e);   // quad_perm [2,3,0,1]
    return e;
  }
}

// ---------------------------------------------------------------------------
// k_latent. Lanes of a wave are split into groups of L = 2^lshift lanes; a
// group owns one voxel and lane ql of the group owns the subject quads
// q = ql, ql + L, ... (4 subjects per Philox block, one float4 store). Each
// wave walks `viter` voxel groups. XREG keeps the lane's covariates in
// registers (one quad per lane, N <= 4 L). FUSED also forms
// m_j = X'(Z_j - b WM_j) from the fresh draws (no re-read of Z).
// ---------------------------------------------------------------------------
typedef float f4v __attribute__((ext_vector_type(4)));

// ZS: 0 = store Z, 1 = store Z with nontemporal (streaming) stores,
// 2 = do not keep Z (only with FUSED: X'Z is formed from the draws in registers)
// the body of block blk of chain `chain` (sweep: the Philox sweep word)
template <int P, bool XREG, bool FUSED, int ZS>
__device__ __forceinline__ void latent_body(const DevProblem &pb, int lshift, int reps, int viter, uint32_t sweep,
                                            int chain, int blk)
{
  static_assert(ZS != 2 || FUSED, "Z can only be dropped when X'Z is fused");
  const int lane = threadIdx.x & 63;
  const int wave = blk * 4 + (threadIdx.x >> 6);
  const int vpw = 64 >> lshift;
  const int sub = lane >> lshift;
  const int ql = lane & ((1 << lshift) - 1);
  const float b = pb.bwm[chain];
  const float eps = pb.trunc_eps;
  float4 xr[P];
  if (XREG) {
    const int i0 = 4 * min(ql, pb.NQ - 1);
#pragma unroll
    for (int c = 0; c < P; c++) xr[c] = *reinterpret_cast<const float4 *>(pb.X + (size_t)c * pb.NS + i0);
  }
  // per-voxel inputs of the next iteration are loaded one iteration ahead
  const float *betc = pb.beta + (size_t)chain * pb.V * P;
  auto voxel_of = [&](int it) { return pb.vb + (wave * viter + it) * vpw + sub; };
  int sn = min(voxel_of(0), pb.ve - 1);
  float bs_n[P];
#pragma unroll
  for (int c = 0; c < P; c++) bs_n[c] = betc[(size_t)sn * P + c];
  float wm_n = pb.wm[sn];
  uint32_t y_n = XREG ? pb.ybits[(size_t)sn * pb.NW + (ql >> 3)] : 0u;
  for (int it = 0; it < viter; it++) {
    const int grp = wave * viter + it;
    if (pb.vb + grp * vpw >= pb.ve) break;  // wave-uniform
    const int s = pb.vb + grp * vpw + sub;
    const bool valid = s < pb.ve;
    const int sv = valid ? s : pb.ve - 1;
    float bs[P];
#pragma unroll
    for (int c = 0; c < P; c++) bs[c] = bs_n[c];
    const float bw = b * wm_n;
    const uint32_t y0 = y_n;
    if (it + 1 < viter) {
      sn = min(voxel_of(it + 1), pb.ve - 1);
#pragma unroll
      for (int c = 0; c < P; c++) bs_n[c] = betc[(size_t)sn * P + c];
      wm_n = pb.wm[sn];
      if (XREG) y_n = pb.ybits[(size_t)sn * pb.NW + (ql >> 3)];
    }
    float *zrow = pb.Z + ((size_t)chain * pb.V + sv) * pb.NS;
    const uint32_t *yrow = pb.ybits + (size_t)sv * pb.NW;
    float acc[P];
#pragma unroll
    for (int c = 0; c < P; c++) acc[c] = 0.0f;
    for (int r = 0; r < reps; r++) {
      const int q = ql + (r << lshift);
      if (valid && q < pb.NQ) {
        const int i0 = 4 * q;
        float4 x[P];
#pragma unroll
        for (int c = 0; c < P; c++)
          x[c] = XREG ? xr[c] : *reinterpret_cast<const float4 *>(pb.X + (size_t)c * pb.NS + i0);
        float mu[4] = {bw, bw, bw, bw};
#pragma unroll
        for (int c = 0; c < P; c++) {
          mu[0] = fmaf(bs[c], x[c].x, mu[0]);
          mu[1] = fmaf(bs[c], x[c].y, mu[1]);
          mu[2] = fmaf(bs[c], x[c].z, mu[2]);
          mu[3] = fmaf(bs[c], x[c].w, mu[3]);
        }
        const uint32_t yw = (XREG ? y0 : yrow[q >> 3]) >> ((q & 7) * 4);
        uint32_t rr[4];
        philox10((uint32_t)s * (uint32_t)pb.NQ + (uint32_t)q, sweep, TAG_LATENT, (uint32_t)chain,
                 pb.key0, pb.key1, rr);
        float z[4];
        latent_quad(mu, yw, rr, eps, z);
        if (i0 + 3 >= pb.N) {  // last quad of a row: zero the padding subjects
#pragma unroll
          for (int l = 1; l < 4; l++) z[l] = (i0 + l < pb.N) ? z[l] : 0.0f;
        }
        if (ZS == 0) {
          *reinterpret_cast<float4 *>(zrow + i0) = make_float4(z[0], z[1], z[2], z[3]);
        } else if (ZS == 1) {
          const f4v zv = {z[0], z[1], z[2], z[3]};
          __builtin_nontemporal_store(zv, reinterpret_cast<f4v *>(zrow + i0));
        }
        if (FUSED) {
#pragma unroll
          for (int c = 0; c < P; c++) {
            acc[c] = fmaf(x[c].x, z[0], acc[c]);
            acc[c] = fmaf(x[c].y, z[1], acc[c]);
            acc[c] = fmaf(x[c].z, z[2], acc[c]);
            acc[c] = fmaf(x[c].w, z[3], acc[c]);
          }
        }
      }
    }
    if (FUSED) {
      if (lshift == 6) {
        const float tot = wave_reduce_scatter<P>(acc, lane);
        const int c = red_cov<P>(lane);
        if (c >= 0) pb.M[((size_t)chain * pb.V + s) * P + c] = tot;
      } else {
        for (int off = (1 << lshift) >> 1; off > 0; off >>= 1) {
#pragma unroll
          for (int c = 0; c < P; c++) acc[c] += __shfl_xor(acc[c], off, 64);
        }
        if (valid && ql == 0) {
          float *m = pb.M + ((size_t)chain * pb.V + s) * P;
#pragma unroll
          for (int c = 0; c < P; c++) m[c] = acc[c];
        }
      }
    }
  }
}

template <int P, bool XREG, bool FUSED, int ZS>
__global__ __launch_bounds__(256) void k_latent(DevProblem pb, int lshift, int reps, int viter, uint32_t sweep)
{
  latent_body<P, XREG, FUSED, ZS>(pb, lshift, reps, viter, sweep + *pb.sweep_base, blockIdx.y, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Exact single draw by inversion (the arithmetic of latent_quad for one
// subject): W ~ N(0,1) truncated below at a = eps - s mu, Z = mu + s W.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float latent_one(float mu, bool y, float u, float eps)
{
  const float s = y ? 1.0f : -1.0f;
  const float al = fmaf(-s, mu, eps);
  const float qt = q_tail(fabsf(al));
  const float uq = u * qt;
  const float pn = u - uq;
  const bool neg = al < 0.0f;
  const bool flip = neg && pn > 0.5f;
  const float arg = neg ? (flip ? (1.0f - u) + uq : pn) : uq;
  float w = q_inv(arg);
  w = flip ? -w : w;
  w = (w < al) ? al : w;
  if (al >= 8.0f) w = fmaxf(tail_w(al, u), al);
  return fmaf(s, w, mu);
}

// ---------------------------------------------------------------------------
// k_latent_vl: fused latent draw + X'Z with voxels across lanes, for the form
// that does not keep Z. A wave holds 64/G voxels; the G lanes of a voxel take
// G contiguous ranges of its subject quads. X is staged once per block in LDS
// as [q][c] float4 and read at addresses shared by all lanes of a subject
// group (broadcast), so the per-voxel inputs live in registers, the covariate
// products need no cross-lane traffic, and the G partial sums of m_j meet in
// log2(G) lane exchanges once per voxel. Blocks walk voxel tiles of
// 4 * 64/G voxels, grid-strided.
//
// Draws: BMF = false draws every subject by inversion (latent_quad, the
// draws of k_latent bit for bit). BMF = true is accept-or-fallback, also
// exact in distribution. The four uniforms of the
// quad's Philox block give two Box-Muller pairs (R = sqrt(-2 ln u0),
// W0 = R cos 2 pi u1, W1 = R sin 2 pi u1; likewise u2, u3), hardware
// v_log/v_sin/v_cos (1e-7 relative / absolute). Z = mu + W is kept when it
// falls in the interval of y (eps, 500) or (-500, -eps). Otherwise the draw
// is queued in the lane's own LDS list and later redrawn exactly, by
// inversion, from an independent Philox block (tag FALLBACK, sub = subject
// in quad). An accepted candidate is a TN draw and a rejected one is
// replaced by an independent TN draw, so every Z is TN. About 89 % of draws
// take the cheap path at the benchmark's data (y = 0 draws reject when
// mu + W > -eps, y = 1 draws mostly). A drain spreads the wave's queued
// draws over all 64 lanes (each lane publishes its voxel inputs; a lane
// finds an item's owner by binary search over the wave's queue offsets) and
// writes each redraw back to the owner's slot; every owner then adds its own
// redraws in queue order, so the sums are deterministic.
// ---------------------------------------------------------------------------
constexpr int VL_QUEUE = 16;  // fallback slots per lane: subject index | y << 15

// LDS bytes past X of the accept-or-fallback form: per thread the queue
// (uint16), the redraws (float), its voxel inputs (P coefficients, b WM_j,
// counter base) and its queue offset in the wave.
__host__ __device__ constexpr size_t vl_fallback_lds(int P)
{
  return 256 * (VL_QUEUE * sizeof(uint16_t) + VL_QUEUE * sizeof(float) + (P + 2) * sizeof(float) + sizeof(int));
}

// One block's share: tiles t0, t0 + tstep, ... of chain `chain`; G (1, 2, 4
// or 8) lanes per voxel; WMU writes its FP64 partial to bstat slot `slot`.
template <int P, bool WMU, bool BMF>
__device__ __forceinline__ void latent_vl_body(const DevProblem &pb, int G, uint32_t sweep, int chain, int t0,
                                               int tstep, int slot)
{
  extern __shared__ float4 xs[];  // [NQ][P], then (BMF) the fallback areas
  uint16_t *const qall = reinterpret_cast<uint16_t *>(xs + (size_t)pb.NQ * P);  // [256][VL_QUEUE]
  float *const rall = reinterpret_cast<float *>(qall + 256 * VL_QUEUE);        // [256][VL_QUEUE]
  float *const info = rall + 256 * VL_QUEUE;                                   // [256][P + 1]
  uint32_t *const cbs = reinterpret_cast<uint32_t *>(info + 256 * (P + 1));    // [256] counter bases
  int *const offs = reinterpret_cast<int *>(cbs + 256);                         // [256]
  uint16_t *const queue = qall + threadIdx.x * VL_QUEUE;
  const float *xsf = reinterpret_cast<const float *>(xs);  // x_c,i = xsf[(q P + c) 4 + l]
  const int lgg = G >= 8 ? 3 : G >= 4 ? 2 : G >= 2 ? 1 : 0;
  const int VPW = 64 >> lgg;
  const int lane = threadIdx.x & 63;
  const int vl = lane & (VPW - 1);
  const int g = lane >> (6 - lgg);
  for (int i = threadIdx.x; i < pb.NQ * P; i += blockDim.x) {
    const int q = i / P, c = i - q * P;
    xs[i] = *reinterpret_cast<const float4 *>(pb.X + (size_t)c * pb.NS + 4 * q);
  }
  __syncthreads();
  const float b = pb.bwm[chain];
  const float eps = pb.trunc_eps;
  float xsum[P];
#pragma unroll
  for (int c = 0; c < P; c++) xsum[c] = pb.xsum[c];
  const int R = (pb.NQ + (1 << lgg) - 1) >> lgg;
  const int q0 = g * R;
  const int q1 = min(q0 + R, pb.NQ);
  const int tiles = (pb.ve - pb.vb + 4 * VPW - 1) / (4 * VPW);
  double bpart = 0.0;  // WMU: this thread's share of sum_j WM_j (sum_i Z_ij - X'1 b_j)
  for (int t = t0; t < tiles; t += tstep) {
    const int s = pb.vb + (t * 4 + (threadIdx.x >> 6)) * VPW + vl;
    const bool valid = s < pb.ve;
    const int sv = valid ? s : pb.ve - 1;
    const float wmv = pb.wm[sv];
    const float bw = b * wmv;
    const float *bp = pb.beta + ((size_t)chain * pb.V + sv) * P;
    float bs[P];
#pragma unroll
    for (int c = 0; c < P; c++) bs[c] = bp[c];
    const uint32_t *yrow = pb.ybits + (size_t)sv * pb.NW;
    const uint32_t cbase = (uint32_t)sv * (uint32_t)pb.NQ;
    float acc[P];
#pragma unroll
    for (int c = 0; c < P; c++) acc[c] = 0.0f;
    float zs = 0.0f;
    int nq = 0;  // queued fallback draws of this lane
    // exact redraw of every queued subject, lanes in step, each its own queue
    // redraw every queued subject of the wave, all lanes sharing the work
    auto drain = [&]() {
      const int wb = threadIdx.x & ~63;
      float *mine = info + threadIdx.x * (P + 1);
#pragma unroll
      for (int c = 0; c < P; c++) mine[c] = bs[c];
      mine[P] = bw;
      cbs[threadIdx.x] = cbase;
      int v = nq;  // inclusive scan of the queue lengths over the wave
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
      }
      const int total = __shfl(v, 63, 64);
      offs[threadIdx.x] = v - nq;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int base = 0; base < total; base += 64) {
        const int k = base + lane;
        if (k < total) {
          int o = 0;  // owner: the last lane whose offset is <= k
#pragma unroll
          for (int st = 32; st >= 1; st >>= 1) o += (offs[wb + o + st] <= k) ? st : 0;
          const int ow = wb + o, t = k - offs[ow];
          const uint32_t e = qall[ow * VL_QUEUE + t];
          const int i = e & 0x7FFF, q = i >> 2, l = i & 3;
          const float *oi = info + ow * (P + 1);
          float mu = oi[P];
#pragma unroll
          for (int c = 0; c < P; c++) mu = fmaf(oi[c], xsf[(q * P + c) * 4 + l], mu);
          uint32_t rr[4];
          philox10(cbs[ow] + (uint32_t)q, sweep, TAG_FALLBACK | ((uint32_t)l << 4),
                   (uint32_t)chain, pb.key0, pb.key1, rr);
          rall[ow * VL_QUEUE + t] = latent_one(mu, (e >> 15) != 0, u01f(rr[0]), eps);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int t = 0; __ballot(t < nq) != 0; t++) {
        if (t < nq) {
          const int i = queue[t] & 0x7FFF, q = i >> 2, l = i & 3;
          const float z = rall[threadIdx.x * VL_QUEUE + t];
#pragma unroll
          for (int c = 0; c < P; c++) acc[c] = fmaf(xsf[(q * P + c) * 4 + l], z, acc[c]);
          if (WMU) zs += z;
        }
      }
      __builtin_amdgcn_wave_barrier();
      nq = 0;
    };
    // the trip count is the same for every lane (R), so a drain runs with the
    // whole wave; lanes past their own range skip the body
    for (int r = 0; r < R; r++) {
      const int q = q0 + r;
      if (BMF && __ballot(nq > VL_QUEUE - 4) != 0) drain();  // room for this quad's four
      if (q >= q1) continue;
      float4 x[P];
#pragma unroll
      for (int c = 0; c < P; c++) x[c] = xs[q * P + c];
      float mu[4] = {bw, bw, bw, bw};
#pragma unroll
      for (int c = 0; c < P; c++) {
        mu[0] = fmaf(bs[c], x[c].x, mu[0]);
        mu[1] = fmaf(bs[c], x[c].y, mu[1]);
        mu[2] = fmaf(bs[c], x[c].z, mu[2]);
        mu[3] = fmaf(bs[c], x[c].w, mu[3]);
      }
      const uint32_t yw = yrow[q >> 3] >> ((q & 7) * 4);
      uint32_t rr[4];
      philox10(cbase + (uint32_t)q, sweep, TAG_LATENT, (uint32_t)chain, pb.key0, pb.key1, rr);
      float z[4];
      const int i0 = 4 * q;
      if (!BMF) {
        latent_quad(mu, yw, rr, eps, z);  // X is zero past N: padding subjects add nothing to m_j
        if (WMU && i0 + 3 >= pb.N) {
#pragma unroll
          for (int l = 1; l < 4; l++) z[l] = (i0 + l < pb.N) ? z[l] : 0.0f;
        }
      } else {
      float w[4];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        // R = sqrt(-2 ln u) = sqrt(-2 ln2 log2 u); angle 2 pi u in revolutions for v_sin/v_cos
        const float rad = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u01f(rr[2 * h])));
        const float ang = u01f(rr[2 * h + 1]);
        w[2 * h] = rad * __builtin_amdgcn_cosf(ang);
        w[2 * h + 1] = rad * __builtin_amdgcn_sinf(ang);
      }
#pragma unroll
      for (int l = 0; l < 4; l++) {
        const float zc = mu[l] + w[l];
        const bool yl = (yw >> l) & 1u;
        const bool ok = yl ? (zc > eps && zc < 500.0f) : (zc < -eps && zc > -500.0f);
        const bool live = i0 + l < pb.N;  // padding subjects are neither kept nor queued
        z[l] = (ok && live) ? zc : 0.0f;
        queue[nq] = (uint16_t)((i0 + l) | (yl ? 0x8000 : 0));  // kept only if the count advances
        nq += (!ok && live) ? 1 : 0;
      }
      }
#pragma unroll
      for (int c = 0; c < P; c++) {
        acc[c] = fmaf(x[c].x, z[0], acc[c]);
        acc[c] = fmaf(x[c].y, z[1], acc[c]);
        acc[c] = fmaf(x[c].z, z[2], acc[c]);
        acc[c] = fmaf(x[c].w, z[3], acc[c]);
      }
      if (WMU) zs += (z[0] + z[1]) + (z[2] + z[3]);
    }
    if (BMF) drain();
#pragma unroll
    for (int off = VPW; off < 64; off <<= 1) {
#pragma unroll
      for (int c = 0; c < P; c++) acc[c] += __shfl_xor(acc[c], off, 64);
      if (WMU) zs += __shfl_xor(zs, off, 64);
    }
    if (valid && g == 0) {
      float *m = pb.M + ((size_t)chain * pb.V + s) * P;
#pragma unroll
      for (int c = 0; c < P; c++) m[c] = acc[c];
      if (WMU) {
        double r = (double)zs;
#pragma unroll
        for (int c = 0; c < P; c++) r -= (double)xsum[c] * (double)bs[c];
        bpart += (double)wmv * r;
      }
    }
  }
  if (WMU) {
    double bp[1] = {bpart};
    block_sum_store<1>(bp, pb.bstat + (size_t)chain * BSTAT_SLOTS + slot);
  }
}

// ---------------------------------------------------------------------------
// latent_vl_mfm_body: the inversion path at G = 4 (16 voxels x 4 subject
// groups per wave) with the covariate means on the matrix cores. Per round
// lane l (voxel j = l % 16, group g = l / 16) draws one quad of its group;
// one v_mfma_f32_16x16x4_f32 (FP32 products and sums) forms all 256 means of
// the wave's round. For that shape lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and receives D[4 (l / 16) + v][l % 16], v = 0..3, so with
//   A[4g + v][c] = x_c of subject v of group g's quad,  B[c][j] = beta_c,j,
//   C = b WM_j
// lane (j, g) receives the four means of its own quad. P = 5..8 takes a second
// MFMA for covariates 4..7. An FP32 MFMA holds the SIMD for 32 cycles and
// does not overlap the VALU here, so it pays only when it replaces more FMAs
// than that: for P >= 3 (P x 4 FMAs of the VALU form); P <= 2 keeps the FMA
// chain. X'Z stays on FP32 FMAs (the X'Z product has P of the MFMA's 16 rows
// and would cost more MFMA cycles than the FMAs it replaces). X is staged for
// 4R quads, zero past NQ, so every lane reads in bounds in every round; the
// means differ from the FMA chain of latent_vl_body only in FP32 rounding.
// BX (X'Z as BF16 hi + lo products on the matrix cores): per round one
// v_mfma_f32_16x16x16_bf16 pair forms D += (Z_hi + Z_lo) [X_hi | X_lo], with
// A = Z (rows: the wave's 16 voxels; K: the round's 16 subjects, lane (j, g)
// holding subjects 4g..4g+3 of its quad as it drew them) and B = the BF16 rows
// of X (pb.xbf: column c < P is X_hi of covariate c, P + c its X_lo; lane
// (c, g) loads the same four subjects). A and B use the same (lane, slot) -> k
// map, so the contraction holds whatever order the hardware sums k in. D
// lands as lane (c, g) -> D[4g + i][c], and m_j,c = D[j][c] + D[j][P + c].
// Z never leaves registers; the four cross products of the hi and lo parts
// are all formed, so the error is that of the splits (|v - hi - lo| <=
// 2^-18 |v| per operand) plus FP32 accumulation.
// ---------------------------------------------------------------------------
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// two floats to packed BF16 (round to nearest even), low half first
__device__ __forceinline__ uint32_t bf16_pk(float a, float b)
{
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  const f32x2 v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// hi = bf16(v), lo = bf16(v - hi) of four floats, packed as MFMA operands
__device__ __forceinline__ void split4_bf16(const float (&v)[4], s16x4 &hi, s16x4 &lo)
{
  const uint32_t h0 = bf16_pk(v[0], v[1]), h1 = bf16_pk(v[2], v[3]);
  const uint32_t l0 = bf16_pk(v[0] - __uint_as_float(h0 << 16), v[1] - __uint_as_float(h0 & 0xffff0000u));
  const uint32_t l1 = bf16_pk(v[2] - __uint_as_float(h1 << 16), v[3] - __uint_as_float(h1 & 0xffff0000u));
  hi = __builtin_bit_cast(s16x4, make_uint2(h0, h1));
  lo = __builtin_bit_cast(s16x4, make_uint2(l0, l1));
}

// ---------------------------------------------------------------------------
// Tagged words of the persistent sweep (PTags). Writers store (tag << 32 |
// payload) with one 64-bit agent-scope store; readers load with agent-scope
// loads until the word shows the tag they need. Tag and payload arrive
// together, so no fence orders anything. A wait is bounded: past pb.lam_spin
// polls (0: PT_SPIN_MAX), or once another wait has given up (flags[2]), it
// sets flags[2] (the host reports it) and goes on, so every block ends.
// ---------------------------------------------------------------------------
constexpr int PT_SPIN_MAX = 1 << 20;
constexpr uint32_t PT_DIAG_RECS = 256;  // waits past their bound recorded per launch
__device__ __forceinline__ uint32_t ptag(const PTags &tg, int k)
{
  return (tg.epoch << 16) | (uint32_t)(k + 1);
}
__device__ __forceinline__ unsigned long long tg_ld(const unsigned long long *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, TAG_SCOPE);
}
__device__ __forceinline__ void tg_st(unsigned long long *p, uint32_t tag, uint32_t bits)
{
  __hip_atomic_store(p, ((unsigned long long)tag << 32) | bits, __ATOMIC_RELAXED, TAG_SCOPE);
}
__device__ __forceinline__ void tg_std(unsigned long long *p, uint32_t tag, double v)
{
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  tg_st(p, tag, (uint32_t)(b >> 32));
  tg_st(p + 1, tag, (uint32_t)b);
}
__device__ __forceinline__ double tg_hilo(uint32_t hi, uint32_t lo)
{
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ bool tg_abort(const DevProblem &pb)
{
  return __hip_atomic_load(pb.flags + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}
// payloads of words p[0 .. N), every one showing `tag`. While some lane of
// the wave still misses its tag, the wave polls one word only (the first
// missing word of the first such lane: one cache line per poll, so that many
// waiting waves do not crowd out the producers' memory traffic), then reads
// every lane's words again.
template <int N>
__device__ __forceinline__ void tg_wait(const unsigned long long *p, uint32_t tag, const DevProblem &pb,
                                        uint32_t (&out)[N])
{
  unsigned long long v[N];
  bool ok = true;
#pragma unroll
  for (int c = 0; c < N; c++) {
    v[c] = tg_ld(p + c);
    ok = ok && (uint32_t)(v[c] >> 32) == tag;
  }
  if (__ballot(!ok) != 0ull) {
    const int lim = pb.lam_spin > 0 ? pb.lam_spin : PT_SPIN_MAX;
    int it = 0;
    bool gave_up = false;
    while (__ballot(!ok) != 0ull) {
      // the representative word: first missing word of the first missing lane
      int cm = 0;
      while (cm < N - 1 && (uint32_t)(v[cm] >> 32) == tag) cm++;
      const unsigned long long mine = ok ? 0ull : (unsigned long long)(p + cm);
      const int src = __ffsll((long long)__ballot(!ok)) - 1;
      const unsigned long long *rep = (const unsigned long long *)__shfl(mine, src, 64);
      for (;; it++) {
        const bool past = it >= lim;
        if (past || (pb.lam_spin == 0 && (it & 31) == 31 && tg_abort(pb))) {
          gave_up = true;
          break;
        }
        __builtin_amdgcn_s_sleep(16);
        if ((uint32_t)(tg_ld(rep) >> 32) == tag) break;
      }
      ok = true;
#pragma unroll
      for (int c = 0; c < N; c++) {
        v[c] = tg_ld(p + c);
        ok = ok && (uint32_t)(v[c] >> 32) == tag;
      }
      if (gave_up) break;
    }
    if (gave_up && !ok) {
      const unsigned long long curi = pb.pdiag ? pb.pdiag[4 + blockIdx.x] : 0ull;
      // one record per item of a block: the first of its waits past the bound
      if (it >= lim && pb.pdiag && atomicExch(pb.pdiag + 4 + gridDim.x + blockIdx.x, curi) != curi) {
        const uint32_t slot = atomicAdd(&pb.flags[3], 1u);
        if (slot < PT_DIAG_RECS) {
          int c = 0;  // the first word that does not show the tag
          while (c < N - 1 && (uint32_t)(v[c] >> 32) == tag) c++;
          unsigned long long *r = pb.pdiag + 4 + 2 * gridDim.x + 4 * slot;
          r[0] = (unsigned long long)(p + c);
          r[1] = tag;
          r[2] = v[c];
          r[3] = curi;  // the item the block runs
        }
      }
      atomicOr(&pb.flags[2], 1u);
    }
  }
#pragma unroll
  for (int c = 0; c < N; c++) out[c] = (uint32_t)v[c];
}
template <int N>
__device__ __forceinline__ void tg_waitf(const unsigned long long *p, uint32_t tag, const DevProblem &pb,
                                         float (&out)[N])
{
  uint32_t w[N];
  tg_wait<N>(p, tag, pb, w);
#pragma unroll
  for (int c = 0; c < N; c++) out[c] = __uint_as_float(w[c]);
}

// X of the G = 4 inversion path in LDS: [4R][P] quads, zero past NQ
template <int P>
__device__ __forceinline__ void stage_x_quads(const DevProblem &pb, float4 *xs)
{
  const int R = (pb.NQ + 3) >> 2;
  for (int i = threadIdx.x; i < 4 * R * P; i += blockDim.x) {
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (i < pb.NQ * P) {
      const int q = i / P, c = i - q * P;
      v = *reinterpret_cast<const float4 *>(pb.X + (size_t)c * pb.NS + 4 * q);
    }
    xs[i] = v;
  }
}

// TG (the persistent sweep, relative sweep pk): X is already in LDS; b of
// relative sweep pk - 1 comes from the tagged words (pk > 0; the launch's
// first sweep reads pb.beta, written before the launch); m_j goes to the
// tagged words and, in the launch's last sweep (fin), to pb.M; a piece's
// partial sums go to the tagged words only. (Plain stores of one address
// from blocks of different XCDs in one launch reach memory in no fixed
// order, since each XCD's L2 writes back its own lines: so the plain arrays
// take one store per launch, the last sweep's.)
template <int P, bool WMU, bool TG = false, bool BX = false>
__device__ __forceinline__ void latent_vl_mfm_body(const DevProblem &pb, uint32_t sweep, int chain, int vb, int ve,
                                                   int t0, int tstep, int slot, int ra = 0, int rb = -1,
                                                   float *part = nullptr, const PTags *tg = nullptr, int pk = 0,
                                                   unsigned long long *tpart = nullptr, bool fin = true)
{
  static_assert(P <= 8, "means take at most two MFMAs");
  static_assert(!(TG && WMU), "the persistent sweep keeps b fixed");
  static_assert(!(BX && (TG || WMU)), "the BF16 X'Z form keeps b fixed and runs outside the persistent sweep");
  constexpr bool MFM = P >= 3;
  extern __shared__ float4 xs[];  // [4R][P]
  const int R = (pb.NQ + 3) >> 2;
  if constexpr (!TG) {
    stage_x_quads<P>(pb, xs);
    __syncthreads();
  }
  const float *xsf = reinterpret_cast<const float *>(xs);
  const int lane = threadIdx.x & 63, vl = lane & 15, g = lane >> 4;
  const float b = pb.bwm[chain];
  const float eps = pb.trunc_eps;
  float xsum[P];
#pragma unroll
  for (int c = 0; c < P; c++) xsum[c] = pb.xsum[c];
  // round-0 LDS offsets (floats) of this lane's A entries of the means;
  // covariates past P read a valid X entry against a zero B
  const int ga = vl >> 2, ea = vl & 3;
  const int am0 = ((ga * R) * P + min(g, P - 1)) * 4 + ea;
  const int am1 = ((ga * R) * P + min(g + 4, P - 1)) * 4 + ea;
  const int q0 = g * R;
  const int tiles = (ve - vb + 63) / 64;
  double bpart = 0.0;
  // BX: this lane's B column of the X'Z product (X_hi of covariate vl, or
  // X_lo of covariate vl - P; none past 2P)
  const bool xon = vl < 2 * P;
  const uint16_t *xrow = BX ? pb.xbf + (size_t)(xon ? vl : 0) * pb.NSB : nullptr;
  for (int t = t0; t < tiles; t += tstep) {
    const int s = vb + (t * 4 + (threadIdx.x >> 6)) * 16 + vl;
    const bool valid = s < ve;
    const int sv = valid ? s : ve - 1;
    const float wmv = pb.wm[sv];
    const float bw = b * wmv;
    const float *bp = pb.beta + ((size_t)chain * pb.V + sv) * P;
    float bs[P];
    if (TG && pk > 0) {
      tg_waitf<P>(tg->bt + ((size_t)((pk - 1) & 1) * pb.V + sv) * P, ptag(*tg, pk - 1), pb, bs);
    } else {
#pragma unroll
      for (int c = 0; c < P; c++) bs[c] = bp[c];
    }
    float bm0 = 0.0f, bm1 = 0.0f;  // B of the means: beta_g and beta_(g+4) of this voxel
#pragma unroll
    for (int c = 0; c < P; c++) {
      bm0 = (g == c) ? bs[c] : bm0;
      if (c >= 4) bm1 = (g + 4 == c) ? bs[c] : bm1;
    }
    const f32x4 bw4 = {bw, bw, bw, bw};
    const uint32_t *yrow = pb.ybits + (size_t)sv * pb.NW;
    const uint32_t cbase = (uint32_t)sv * (uint32_t)pb.NQ;
    float acc[P], zs = 0.0f;
#pragma unroll
    for (int c = 0; c < P; c++) acc[c] = 0.0f;
    f32x4 dx = {0.0f, 0.0f, 0.0f, 0.0f};  // BX: D[4g + i][vl] of this tile
    const int rend = rb < 0 ? R : rb;
    int am = ra * 4 * P;  // step of the means' A offsets
    for (int r = ra; r < rend; r++, am += 4 * P) {
      const int q = q0 + r;  // passes NQ only in the last group, where X reads zero
      float4 x[P];  // read before the draw only for the FMA means: not live through it otherwise
      if (!MFM) {
#pragma unroll
        for (int c = 0; c < P; c++) x[c] = xs[q * P + c];
      }
      float mu[4];
      if (MFM) {
        f32x4 m4 = __builtin_amdgcn_mfma_f32_16x16x4f32(xsf[am0 + am], bm0, bw4, 0, 0, 0);
        if (P > 4) m4 = __builtin_amdgcn_mfma_f32_16x16x4f32(xsf[am1 + am], bm1, m4, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; v++) mu[v] = m4[v];
      } else {
#pragma unroll
        for (int v = 0; v < 4; v++) mu[v] = bw;
#pragma unroll
        for (int c = 0; c < P; c++) {
          mu[0] = fmaf(bs[c], x[c].x, mu[0]);
          mu[1] = fmaf(bs[c], x[c].y, mu[1]);
          mu[2] = fmaf(bs[c], x[c].z, mu[2]);
          mu[3] = fmaf(bs[c], x[c].w, mu[3]);
        }
      }
      const int qy = min(q, pb.NQ - 1);
      const uint32_t yw = yrow[qy >> 3] >> ((qy & 7) * 4);
      // BX: B of this round (X_hi or X_lo of four subjects; zero past NQ and
      // past column 2P), loaded from an address clamped into the row
      uint2 xw = make_uint2(0u, 0u);
      if constexpr (BX) {
        xw = *reinterpret_cast<const uint2 *>(xrow + 4 * qy);
        if (!xon || q >= pb.NQ) xw = make_uint2(0u, 0u);
      }
      uint32_t rr[4];
      philox10(cbase + (uint32_t)q, sweep, TAG_LATENT, (uint32_t)chain, pb.key0, pb.key1, rr);
      float z[4];
      latent_quad(mu, yw, rr, eps, z);  // X is zero past N: padding subjects add nothing to m_j
      if (WMU && 4 * q + 3 >= pb.N) {
#pragma unroll
        for (int l = 0; l < 4; l++) z[l] = (4 * q + l < pb.N) ? z[l] : 0.0f;
      }
      if constexpr (BX) {
        s16x4 zh, zl;
        split4_bf16(z, zh, zl);
        const s16x4 xb = __builtin_bit_cast(s16x4, xw);
        dx = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(zh, xb, dx, 0, 0, 0);
        dx = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(zl, xb, dx, 0, 0, 0);
      } else {
        if (MFM) {
#pragma unroll
          for (int c = 0; c < P; c++) x[c] = xs[q * P + c];
        }
#pragma unroll
        for (int c = 0; c < P; c++) {
          acc[c] = fmaf(x[c].x, z[0], acc[c]);
          acc[c] = fmaf(x[c].y, z[1], acc[c]);
          acc[c] = fmaf(x[c].z, z[2], acc[c]);
          acc[c] = fmaf(x[c].w, z[3], acc[c]);
        }
      }
      if (WMU) zs += (z[0] + z[1]) + (z[2] + z[3]);
    }
    if constexpr (BX) {
      // m of voxel 4g + i, covariate vl: D[.][vl] + D[.][vl + P] (2P <= 16:
      // the partner lane is in the same row of 16)
      const int wv = threadIdx.x >> 6;
      const int sw = vb + (t * 4 + wv) * 16 + 4 * g;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float mv = dx[i] + __shfl_down(dx[i], P, 16);
        if (vl < P) {
          if (part) part[(wv * 16 + 4 * g + i) * P + vl] = mv;
          else if (sw + i < ve) pb.M[((size_t)chain * pb.V + sw + i) * P + vl] = mv;
        }
      }
      continue;
    }
#pragma unroll
    for (int off = 16; off < 64; off <<= 1) {
#pragma unroll
      for (int c = 0; c < P; c++) acc[c] += __shfl_xor(acc[c], off, 64);
      if (WMU) zs += __shfl_xor(zs, off, 64);
    }
    if (TG && tpart) {  // a piece of a split tile, persistent sweep: tagged partial sums
      if (g == 0) {
#pragma unroll
        for (int c = 0; c < P; c++)
          tg_st(tpart + ((threadIdx.x >> 6) * 16 + vl) * P + c, ptag(*tg, pk), __float_as_uint(acc[c]));
      }
    } else if (part) {  // a piece of a split tile: partial sums, tile-local voxel order
      if (g == 0) {
#pragma unroll
        for (int c = 0; c < P; c++) part[((threadIdx.x >> 6) * 16 + vl) * P + c] = acc[c];
      }
    } else if (valid && g == 0) {
      float *m = pb.M + ((size_t)chain * pb.V + s) * P;
      if (!TG || fin) {
#pragma unroll
        for (int c = 0; c < P; c++) m[c] = acc[c];
      }
      if constexpr (TG) {
#pragma unroll
        for (int c = 0; c < P; c++) tg_st(tg->mt + (size_t)s * P + c, ptag(*tg, pk), __float_as_uint(acc[c]));
      }
      if (WMU) {
        double r = (double)zs;
#pragma unroll
        for (int c = 0; c < P; c++) r -= (double)xsum[c] * (double)bs[c];
        bpart += (double)wmv * r;
      }
    }
  }
  if (WMU) {
    double bp[1] = {bpart};
    block_sum_store<1>(bp, pb.bstat + (size_t)chain * BSTAT_SLOTS + slot);
  }
}

template <int P, bool WMU, bool BMF>
__global__ __launch_bounds__(256) void k_latent_vl(DevProblem pb, int G, uint32_t sweep)
{
  sweep += *pb.sweep_base;
  latent_vl_body<P, WMU, BMF>(pb, G, sweep, blockIdx.y, blockIdx.x, gridDim.x, blockIdx.x);
}

template <int P>
__device__ void wishart_block(const DevProblem &pb, int chain, int nparts, const double *part, uint32_t sweep,
                              bool tagged = false, const double *apre = nullptr);

// Colour tiling of the inversion path at G = 4 (one chain). Each colour's
// voxels form 64-voxel tiles from its first sorted voxel (vl_range). A launch
// of T tiles over the chip's C compute units runs ceil(T / C) tiles on some
// units where T / C would be even, so a colour's T mod C leftover tiles are
// split, when few, into K pieces by subject rounds (piece k takes rounds
// [kR/K, (k+1)R/K)), which spread over the whole chip. A launch over colour
// set cs runs the whole tiles of each colour in it, then the pieces; each
// piece writes its tile's 64 partial X'Z sums to mpart. The coefficient
// kernel that next reads such a voxel adds its K partials in piece order
// (vl_tail_sum) and writes m_j back: the kernel boundary orders the writes
// before the reads, so no fence or counter is needed (a device-scope fence
// here writes back and invalidates the L2 of the XCD, which slowed the whole
// launch). The pieces draw exactly the variates of the whole tile (same
// counters); only the order of the X'Z sum differs. Tiles and pieces depend
// on the problem alone, so the whole-range launch, the one-colour launches of
// the overlapped sweep and a batch all form the same sums.
template <int P, bool BX = false>
__device__ __forceinline__ void vl_ctile_block(const DevProblem &pb, uint32_t sweep, int chain, int b, int cs)
{
  const VlRange r0 = vl_range(pb, 0), r1 = vl_range(pb, 1);
  const int m0 = (cs & 1) ? r0.nmain : 0, m1 = (cs & 2) ? r1.nmain : 0;
  if (b < m0 + m1) {
    const bool c0 = b < m0;
    latent_vl_mfm_body<P, false, false, BX>(pb, sweep, chain, c0 ? r0.base : r1.base, c0 ? r0.end : r1.end,
                                            c0 ? b : b - m0, 1 << 30, 0);
    return;
  }
  b -= m0 + m1;
  const int p0 = (cs & 1) ? r0.left * r0.K : 0;
  const VlRange r = b < p0 ? r0 : r1;
  if (b >= p0) b -= p0;
  const int tl = b / r.K, k = b - tl * r.K;
  const int R = (pb.NQ + 3) >> 2;
  latent_vl_mfm_body<P, false, false, BX>(pb, sweep, chain, r.base, r.end, r.nmain + tl, 1 << 30, 0, k * R / r.K,
                                          (k + 1) * R / r.K,
                                          pb.mpart + (size_t)chain * pb.mp_stride + r.moff + ((size_t)tl * r.K + k) * 64 * P);
}

// m_j of sorted voxel s when it lies in a leftover tile of its colour: its K
// partial sums added in piece order (the same order wherever it is formed)
template <int P>
__device__ __forceinline__ bool vl_tail_sum(const DevProblem &pb, int chain, int s, float (&m)[P])
{
  const VlRange r = vl_range(pb, s >= pb.n0 ? 1 : 0);
  const int loc = s - r.base, tl = (loc >> 6) - r.nmain;
  if (tl < 0) return false;
  const float *tp = pb.mpart + (size_t)chain * pb.mp_stride + r.moff + (size_t)tl * r.K * 64 * P + (size_t)(loc & 63) * P;
#pragma unroll
  for (int c = 0; c < P; c++) m[c] = 0.0f;
  for (int kk = 0; kk < r.K; kk++) {
#pragma unroll
    for (int c = 0; c < P; c++) m[c] += tp[(size_t)kk * 64 * P + c];
  }
  return true;
}

// The persistent sweep's latent item b of colour c (the numbering of
// vl_ctile_block over that colour alone: whole tiles, then pieces), relative
// sweep pk, Philox sweep word `sweep`
template <int P>
__device__ __forceinline__ void vl_ctile_block_tg(const DevProblem &pb, const PTags &tg, int pk, uint32_t sweep, int b,
                                                  int c)
{
  const VlRange r = vl_range(pb, c);
  if (b < r.nmain) {
    latent_vl_mfm_body<P, false, true>(pb, sweep, 0, r.base, r.end, b, 1 << 30, 0, 0, -1, nullptr, &tg, pk, nullptr,
                                       pk == tg.K - 1);
    return;
  }
  b -= r.nmain;
  const int tl = b / r.K, k = b - tl * r.K;
  const int R = (pb.NQ + 3) >> 2;
  latent_vl_mfm_body<P, false, true>(pb, sweep, 0, r.base, r.end, r.nmain + tl, 1 << 30, 0, k * R / r.K,
                                     (k + 1) * R / r.K, nullptr, &tg, pk,
                                     tg.mpt + r.moff + ((size_t)tl * r.K + k) * 64 * P);
}

// vl_tail_sum from the tagged partials of relative sweep pk (same order)
template <int P>
__device__ __forceinline__ bool vl_tail_sum_tg(const DevProblem &pb, const PTags &tg, int pk, int s, float (&m)[P])
{
  const VlRange r = vl_range(pb, s >= pb.n0 ? 1 : 0);
  const int loc = s - r.base, tl = (loc >> 6) - r.nmain;
  if (tl < 0) return false;
  const unsigned long long *tp = tg.mpt + r.moff + (size_t)tl * r.K * 64 * P + (size_t)(loc & 63) * P;
#pragma unroll
  for (int c = 0; c < P; c++) m[c] = 0.0f;
  for (int kk = 0; kk < r.K; kk++) {
    float v[P];
    tg_waitf<P>(tp + (size_t)kk * 64 * P, ptag(tg, pk), pb, v);
#pragma unroll
    for (int c = 0; c < P; c++) m[c] += v[c];
  }
  return true;
}

// the inversion path at G = 4; a kernel of its own so its registers are its
// own. cs > 0: a colour-tiled launch over colour set cs; cs = 0: tiles of
// [pb.vb, pb.ve) strided over the grid. With wish = 1, block 0 draws the CAR
// precision of sweep wsweep (the sweep before: its colour-1 partials are
// complete, and nothing here reads Lambda) beside the latent blocks, instead
// of a launch of its own.
template <int P, bool WMU, bool BX = false>
__global__ __launch_bounds__(256, P <= 7 ? 8 : 7) void k_latent_vl_mfm(DevProblem pb, uint32_t sweep, int wish,
                                                                       uint32_t wsweep, int cs)
{
  const uint32_t base = *pb.sweep_base;
  const int chain = blockIdx.y;
  // the draw keeps its FP64 matrices in LDS, so it fits the latent kernel's registers
  if (wish && blockIdx.x == 0) {
    constexpr int NE = P * (P + 1) / 2;
    wishart_block<P>(pb, chain, (pb.V - pb.n0 + 255) / 256, pb.partials + (size_t)chain * pb.max_parts * NE,
                     wsweep + base);
    return;
  }
  const int b = blockIdx.x - wish;
  if (!WMU && cs) vl_ctile_block<P, BX>(pb, sweep + base, chain, b, cs);
  else latent_vl_mfm_body<P, WMU, false, BX>(pb, sweep + base, chain, pb.vb, pb.ve, b, gridDim.x - wish, b);
}

// ---------------------------------------------------------------------------
// WM coefficient b (covarGPU.cu:588-724, GPU semantics: flat prior).
//   S = sum_j WM_j sum_i (Z_ij - x_i'b_j)   (betaMGPU, X'1 . b_j per voxel)
//   W = N sum_j WM_j^2                      (betaVGPU, host FP64 at set-up)
//   b ~ N(S / W, 1 / W)
// k_bstat forms block partials of S from the stored Z (one thread per voxel,
// FP64 accumulation); the Z-free latent kernel forms the same partials from
// its draws. k_bwm sums the partials in slot order and draws b.
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(256) void k_bstat(DevProblem pb)
{
  const int chain = blockIdx.y;
  const float *betc = pb.beta + (size_t)chain * pb.V * P;
  double part = 0.0;
  for (int s = blockIdx.x * 256 + threadIdx.x; s < pb.V; s += gridDim.x * 256) {
    const float4 *zrow = reinterpret_cast<const float4 *>(pb.Z + ((size_t)chain * pb.V + s) * pb.NS);
    float zs = 0.0f;
    for (int q = 0; q < pb.NQ; q++) {
      const float4 z = zrow[q];
      const int i0 = 4 * q;
      zs += z.x + ((i0 + 1 < pb.N) ? z.y : 0.0f) + ((i0 + 2 < pb.N) ? z.z : 0.0f) +
            ((i0 + 3 < pb.N) ? z.w : 0.0f);
    }
    double r = (double)zs;
#pragma unroll
    for (int c = 0; c < P; c++) r -= (double)pb.xsum[c] * (double)betc[(size_t)s * P + c];
    part += (double)pb.wm[s] * r;
  }
  double pp[1] = {part};
  block_sum_store<1>(pp, pb.bstat + (size_t)chain * BSTAT_SLOTS + blockIdx.x);
}

__global__ __launch_bounds__(256) void k_bwm(DevProblem pb, int nparts, uint32_t sweep)
{
  sweep += *pb.sweep_base;
  const int chain = blockIdx.y;
  __shared__ double S[1];
  sum_partials<1>(nparts, pb.bstat + (size_t)chain * BSTAT_SLOTS, S);
  if (threadIdx.x == 0) {
    uint32_t rr[4];
    philox10(0u, sweep, TAG_BWM, (uint32_t)chain, pb.key0, pb.key1, rr);
    const double n = sqrt(-2.0 * log(u01d(rr[0]))) * cospi(2.0 * u01d(rr[1]));
    const double var = 1.0 / pb.wm2;
    pb.bwm[chain] = (float)(S[0] * var + sqrt(var) * n);
  }
}

// ---------------------------------------------------------------------------
// k_xtz_f32: m_j = X'Z_j in FP32 FMA, reading Z once (b WM_j X'1 is taken off
// by the coefficient kernel, with the b of the sweep). Same lane
// grouping and voxel walk as k_latent (covarGPU.cu:327-352 contraction).
// ---------------------------------------------------------------------------
template <int P, bool XREG>
__global__ __launch_bounds__(256) void k_xtz_f32(DevProblem pb, int lshift, int reps, int viter)
{
  const int chain = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int vpw = 64 >> lshift;
  const int sub = lane >> lshift;
  const int ql = lane & ((1 << lshift) - 1);
  float4 xr[P];
  if (XREG) {
    const int i0 = 4 * min(ql, pb.NQ - 1);
#pragma unroll
    for (int c = 0; c < P; c++) xr[c] = *reinterpret_cast<const float4 *>(pb.X + (size_t)c * pb.NS + i0);
  }
  for (int it = 0; it < viter; it++) {
    const int grp = wave * viter + it;
    if (pb.vb + grp * vpw >= pb.ve) break;  // wave-uniform
    const int s = pb.vb + grp * vpw + sub;
    const bool valid = s < pb.ve;
    float acc[P];
#pragma unroll
    for (int c = 0; c < P; c++) acc[c] = 0.0f;
    if (valid) {
      const float *zrow = pb.Z + ((size_t)chain * pb.V + s) * pb.NS;
      for (int r = 0; r < reps; r++) {
        const int q = ql + (r << lshift);
        if (q >= pb.NQ) break;
        const int i0 = 4 * q;
        const float4 z = *reinterpret_cast<const float4 *>(zrow + i0);
#pragma unroll
        for (int c = 0; c < P; c++) {
          const float4 x = XREG ? xr[c] : *reinterpret_cast<const float4 *>(pb.X + (size_t)c * pb.NS + i0);
          acc[c] = fmaf(x.x, z.x, acc[c]);
          acc[c] = fmaf(x.y, z.y, acc[c]);
          acc[c] = fmaf(x.z, z.z, acc[c]);
          acc[c] = fmaf(x.w, z.w, acc[c]);
        }
      }
    }
    for (int off = (1 << lshift) >> 1; off > 0; off >>= 1) {
#pragma unroll
      for (int c = 0; c < P; c++) acc[c] += __shfl_xor(acc[c], off, 64);
    }
    if (valid && ql == 0) {
      float *m = pb.M + ((size_t)chain * pb.V + s) * P;
#pragma unroll
      for (int c = 0; c < P; c++) m[c] = acc[c];
    }
  }
}

// ---------------------------------------------------------------------------
// k_xtz_mfma: X'Z on the matrix cores, v_mfma_f32_16x16x32_bf16. Z and X are
// split into hi + lo BF16 parts (hi = bf16(v), lo = bf16(v - hi); |v - hi - lo|
// <= 2^-18 |v|); FP32 accumulation. A wave owns 16 voxels (the MFMA rows) and
// walks the subjects 32 at a time (K). B holds [X_hi | X_lo] in its first 2P
// columns, so D = (Z_hi + Z_lo)[X_hi | X_lo] from two MFMAs per block and
// m_j,c = D[j][c] + D[j][P + c]. A and B use the same (lane, slot) -> k map,
// so the contraction is right whatever order the hardware sums k in; the
// output layout is lane l -> D[4 (l / 16) + i][l % 16]. X is constant, so its
// split is formed once (k_xbf, pb.xbf) and each lane loads its eight B values
// as 16 bytes; Z is the stream, and each lane loads the next K step's eight
// values before it splits and multiplies the current ones.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void split8(const float (&v)[8], bf16x8 &hi, bf16x8 &lo)
{
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const __bf16 h = (__bf16)v[j];
    hi[j] = h;
    lo[j] = (__bf16)(v[j] - (float)h);
  }
}

// eight consecutive floats of a row of length NS (a multiple of 4) from
// element k; float4 halves past the row end read as zero
__device__ __forceinline__ void load8(const float *row, int k, int NS, float (&v)[8])
{
  float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
  if (k < NS) a = *reinterpret_cast<const float4 *>(row + k);
  if (k + 4 < NS) b = *reinterpret_cast<const float4 *>(row + k + 4);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

// BF16 hi/lo rows of X: xbf[c][k] = hi(X[c][k]) for c < P, lo(X[c - P][k])
// for P <= c < 2P, with the same conversions k_xtz_mfma made in its loop
__global__ __launch_bounds__(256) void k_xbf(const float *X, int P, int NS, int NSB, uint16_t *xbf)
{
  const int n = 2 * P * NSB;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const int c = i / NSB, k = i - c * NSB;
    const float v = k < NS ? X[(size_t)(c < P ? c : c - P) * NS + k] : 0.0f;
    const __bf16 h = (__bf16)v;
    const __bf16 o = c < P ? h : (__bf16)(v - (float)h);
    xbf[i] = __builtin_bit_cast(uint16_t, o);
  }
}

hipError_t launch_xbf(const DevProblem &pb, uint16_t *xbf, hipStream_t st)
{
  const int n = 2 * pb.P * pb.NSB;
  hipLaunchKernelGGL(k_xbf, dim3((unsigned)std::max(1, std::min((n + 255) / 256, 1024))), dim3(256), 0, st, pb.X,
                     pb.P, pb.NS, pb.NSB, xbf);
  return hipGetLastError();
}

// load8 without branches: both float4 loads always issue, from addresses
// clamped into the row (NS >= 4, a multiple of 4), and values past the row
// end are zeroed by selects
__device__ __forceinline__ void load8c(const float *row, int k, int NS, float (&v)[8])
{
  float4 a = *reinterpret_cast<const float4 *>(row + min(k, NS - 4));
  float4 b = *reinterpret_cast<const float4 *>(row + min(k + 4, NS - 4));
  const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
  if (k >= NS) a = zero;
  if (k + 4 >= NS) b = zero;
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

// block blk of chain `chain`: four waves of 16 voxels each
template <int P>
__device__ __forceinline__ void xtz_mfma_body(const DevProblem &pb, int chain, int blk)
{
  static_assert(2 * P <= 16, "hi/lo columns must fit 16");
  const int lane = threadIdx.x & 63;
  const int tile = blk * 4 + (threadIdx.x >> 6);
  const int s0 = pb.vb + tile * 16;
  if (s0 >= pb.ve) return;  // wave-uniform
  const int r = lane & 15, g = lane >> 4;
  const int srow = min(s0 + r, pb.ve - 1);
  const float *zrow = pb.Z + ((size_t)chain * pb.V + srow) * pb.NS;
  const int col = r;  // B column of this lane
  const bool xon = col < 2 * P;
  const uint16_t *xrow = pb.xbf + (size_t)(xon ? col : 0) * pb.NSB;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  float zv[8];
  load8c(zrow, 8 * g, pb.NS, zv);
  for (int k0 = 0; k0 < pb.NS; k0 += 32) {
    const int k = k0 + 8 * g;
    // B of this step first, then the next step's Z: loads retire in order,
    // so waiting for B leaves the next step's loads in flight. The loads are
    // unconditional (addresses clamped into the row, values past it zeroed
    // by selects), so no branch hides them from the wait counts.
    uint4 xw = *reinterpret_cast<const uint4 *>(xrow + min(k, pb.NSB - 8));
    if (!(xon && k < pb.NSB)) xw = make_uint4(0u, 0u, 0u, 0u);
    float zn[8];
    load8c(zrow, k + 32, pb.NS, zn);  // the next step (zero past the row)
    bf16x8 zh, zl;
    split8(zv, zh, zl);
    const bf16x8 xb = __builtin_bit_cast(bf16x8, xw);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(zh, xb, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(zl, xb, acc, 0, 0, 0);
#pragma unroll
    for (int j = 0; j < 8; j++) zv[j] = zn[j];
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const float other = __shfl_down(acc[i], P, 16);  // D[row][col + P]
    const int s = s0 + 4 * g + i;
    if (col < P && s < pb.ve) pb.M[((size_t)chain * pb.V + s) * P + col] = acc[i] + other;
  }
}

template <int P>
__global__ __launch_bounds__(256) void k_xtz_mfma(DevProblem pb)
{
  xtz_mfma_body<P>(pb, blockIdx.y, blockIdx.x);
}

// ---------------------------------------------------------------------------
// k_wishart / wishart_block: sums the per-block FP64 partials of T in block
// order, forms S = (T+I)^-1 and L = chol(S) in FP64 (covarGPU.cu:1086-1101),
// draws the Bartlett factor A (A_ii^2 ~ chi2(df-i) by Marsaglia-Tsang gamma,
// A_ij ~ N(0,1)) and writes Lambda = (L A)(L A)' (cholesky.cpp:259-281).
// The sweep index (Philox counter word) is passed in by the host. Runs on the
// first wave of a block.
// ---------------------------------------------------------------------------
__device__ double gamma_mt(double a, uint32_t row, uint32_t sweep, uint32_t chain, uint32_t k0,
                           uint32_t k1)
{
  const double d = a - 1.0 / 3.0;
  const double c = 1.0 / sqrt(9.0 * d);
  for (uint32_t t = 0; t < 64; t++) {
    uint32_t rr[4];
    philox10(row, sweep, TAG_GAMMA | (t << 4), chain, k0, k1, rr);
    double x = sqrt(-2.0 * log(u01d(rr[0]))) * cospi(2.0 * u01d(rr[1]));
    double v = 1.0 + c * x;
    if (v <= 0.0) continue;
    v = v * v * v;
    if (log(u01d(rr[2])) < 0.5 * x * x + d - d * v + d * log(v)) return d * v;
  }
  return d;
}

template <int P>
__device__ void chol_lower(double (&a)[P][P])
{
  for (int j = 0; j < P; j++) {
    double d = a[j][j];
    for (int k = 0; k < j; k++) d -= a[j][k] * a[j][k];
    d = sqrt(d);
    a[j][j] = d;
    for (int i = j + 1; i < P; i++) {
      double s = a[i][j];
      for (int k = 0; k < j; k++) s -= a[i][k] * a[j][k];
      a[i][j] = s / d;
    }
    for (int i = 0; i < j; i++) a[i][j] = 0.0;
  }
}

// Entry idx = i P + j of the Bartlett factor of the Wishart draw of sweep
// `sweep`: sqrt(chi2(df - i)) on the diagonal, N(0, 1) below, 0 above. It
// depends on the sweep index alone, so it can be drawn ahead (k_overlap).
template <int P>
__device__ __forceinline__ double bartlett_entry(const DevProblem &pb, int chain, int idx, uint32_t sweep)
{
  const int i = idx / P, j = idx % P;
  double val = 0.0;
  if (i == j) {
    val = sqrt(2.0 * gamma_mt(0.5 * (double)(pb.df - i), (uint32_t)i, sweep, (uint32_t)chain, pb.key0, pb.key1));
  } else if (j < i) {
    uint32_t rr[4];
    philox10((uint32_t)idx, sweep, TAG_WNORM, (uint32_t)chain, pb.key0, pb.key1, rr);
    val = sqrt(-2.0 * log(u01d(rr[0]))) * cospi(2.0 * u01d(rr[1]));
  }
  return val;
}

// ---------------------------------------------------------------------------
// The Wishart algebra spread over the 64 lanes of one wave, matrices in LDS
// (row-major P x P, P at run time up to PW_MAX): used above P = 6. Every
// entry is formed by the operations of the one-lane form in the same order
// (a right-looking Cholesky subtracts the products of an entry in the order
// k = 0, 1, ... as the left-looking loop does; the inverse runs one column
// per lane; the products one entry per lane), so the result does not depend
// on the lane count.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// lower Cholesky factor in place, zeros above the diagonal (chol_lower)
__device__ void chol_lower_wave(double *a, int P, int lane)
{
  for (int j = 0; j < P; j++) {
    if (lane == 0) a[j * P + j] = sqrt(a[j * P + j]);
    wave_sync();
    for (int i = j + 1 + lane; i < P; i += 64) a[i * P + j] = a[i * P + j] / a[j * P + j];
    wave_sync();
    // the trailing entries (i, c), j < c <= i, subtract their term k = j
    const int m = P - j - 1, cnt = m * (m + 1) / 2;
    for (int e = lane; e < cnt; e += 64) {
      int ii = 0;
      while ((ii + 1) * (ii + 2) / 2 <= e) ii++;
      const int i = j + 1 + ii, c = j + 1 + (e - ii * (ii + 1) / 2);
      a[i * P + c] -= a[i * P + j] * a[c * P + j];
    }
    wave_sync();
  }
  for (int e = lane; e < P * P; e += 64)
    if (e % P > e / P) a[e] = 0.0;
  wave_sync();
}

// T packed lower triangle (row-major), A the Bartlett factor; S, inv, B and
// Y (P x P each) are LDS work areas. Writes Lambda (FP32, FP64, and tagged
// when `tagged`) and the pair sums.
__device__ void wishart_algebra_wave(const DevProblem &pb, int chain, int P, const double *T, const double *A,
                                     bool tagged, double *S, double *inv, double *B, double *Y, int lane)
{
  for (int e = lane; e < P * P; e += 64) {
    const int i = e / P, j = e % P;
    const int hi = i > j ? i : j, lo = i > j ? j : i;
    const double t = T[hi * (hi + 1) / 2 + lo];
    S[e] = i == j ? t + 1.0 : t;
    pb.tsum[(size_t)chain * P * P + e] = t;
  }
  wave_sync();
  chol_lower_wave(S, P, lane);
  if (lane < P) {  // column `lane` of the inverse: forward, then back substitution
    const int col = lane;
    double *y = Y + (size_t)col * P;
    for (int i = 0; i < P; i++) {
      double v = (i == col) ? 1.0 : 0.0;
      for (int k = 0; k < i; k++) v -= S[i * P + k] * y[k];
      y[i] = v / S[i * P + i];
    }
    for (int i = P - 1; i >= 0; i--) {
      double v = y[i];
      for (int k = i + 1; k < P; k++) v -= S[k * P + i] * inv[k * P + col];
      inv[i * P + col] = v / S[i * P + i];
    }
  }
  wave_sync();
  chol_lower_wave(inv, P, lane);
  for (int e = lane; e < P * P; e += 64) {
    const int i = e / P, j = e % P;
    double v = 0.0;
    if (j <= i)
      for (int k = j; k <= i; k++) v += inv[i * P + k] * A[k * P + j];
    B[e] = v;
  }
  wave_sync();
  float *lam = pb.lam + (size_t)chain * P * P;
  double *lam64 = pb.lam64 + (size_t)chain * P * P;
  for (int e = lane; e < P * P; e += 64) {
    const int i = e / P, j = e % P;
    const int K = i < j ? i : j;
    double v = 0.0;
    for (int k = 0; k <= K; k++) v += B[i * P + k] * B[j * P + k];
    lam64[e] = v;
    lam[e] = (float)v;
    if (tagged)
      __hip_atomic_store(pb.lamtag + (size_t)chain * P * P + e,
                         ((unsigned long long)(pb.lam_tag ? pb.lam_tag : 1) << 32) |
                             (unsigned long long)__float_as_uint((float)v),
                         __ATOMIC_RELAXED, TAG_SCOPE);
  }
}

// S = (T + I)^-1, its Cholesky factor L, B = L A and Lambda = B B' in FP64
// (cholesky.cpp:245-286 steps), on one lane with the matrices in registers
// (P <= 6); T packed lower triangle.
// plain = false: Lambda goes to the tagged words only (no pb.lam, lam64, tsum)
template <int P>
__device__ __forceinline__ void wishart_algebra(const DevProblem &pb, int chain, const double *T,
                                                const double (&A)[P][P], bool tagged, double (&S)[P][P],
                                                double (&inv)[P][P], double (&B)[P][P], double (&y)[P],
                                                bool plain = true)
{
  for (int a = 0, e = 0; a < P; a++)
    for (int c = 0; c <= a; c++, e++) {
      S[a][c] = T[e];
      S[c][a] = T[e];
      if (plain) {
        pb.tsum[(size_t)chain * P * P + a * P + c] = T[e];
        pb.tsum[(size_t)chain * P * P + c * P + a] = T[e];
      }
    }
  for (int a = 0; a < P; a++) S[a][a] += 1.0;
  chol_lower<P>(S);
  for (int col = 0; col < P; col++) {
    for (int i = 0; i < P; i++) {
      double v = (i == col) ? 1.0 : 0.0;
      for (int k = 0; k < i; k++) v -= S[i][k] * y[k];
      y[i] = v / S[i][i];
    }
    for (int i = P - 1; i >= 0; i--) {
      double v = y[i];
      for (int k = i + 1; k < P; k++) v -= S[k][i] * inv[k][col];
      inv[i][col] = v / S[i][i];
    }
  }
  chol_lower<P>(inv);
  for (int i = 0; i < P; i++)
    for (int j = 0; j < P; j++) {
      double v = 0.0;
      if (j <= i)
        for (int k = j; k <= i; k++) v += inv[i][k] * A[k][j];
      B[i][j] = v;
    }
  float *lam = pb.lam + (size_t)chain * P * P;
  double *lam64 = pb.lam64 + (size_t)chain * P * P;
  for (int i = 0; i < P; i++)
    for (int j = 0; j < P; j++) {
      const int K = i < j ? i : j;
      double v = 0.0;
      for (int k = 0; k <= K; k++) v += B[i][k] * B[j][k];
      if (plain) {
        lam64[i * P + j] = v;
        lam[i * P + j] = (float)v;
      }
      if (tagged)
        __hip_atomic_store(pb.lamtag + (size_t)chain * P * P + i * P + j,
                           ((unsigned long long)(pb.lam_tag ? pb.lam_tag : 1) << 32) |
                               (unsigned long long)__float_as_uint((float)v),
                           __ATOMIC_RELAXED, TAG_SCOPE);
    }
}

// Block of 256 threads: partial sums, Bartlett variates, FP64 algebra.
// tagged: Lambda is also written to lamtag, each entry with tag 1 beside its
// bits, for blocks of the same launch that wait for it (k_overlap).
// apre: the Bartlett factor of this sweep, drawn ahead (else drawn here).
template <int P>
__device__ void wishart_block(const DevProblem &pb, int chain, int nparts, const double *part,
                              uint32_t sweep, bool tagged, const double *apre)
{
  constexpr int NE = P * (P + 1) / 2;
  __shared__ double T[NE];
  __shared__ double A[P][P];
  sum_partials<NE>(nparts, part, T);
  const int lane = threadIdx.x;
  if (lane >= 64) return;
  for (int idx = lane; idx < P * P; idx += 64)
    A[idx / P][idx % P] = apre ? apre[idx] : bartlett_entry<P>(pb, chain, idx, sweep);
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  // the FP64 algebra: on one lane in registers up to P = 6; above, over the
  // wave with its matrices in LDS (few registers, so that the kernels that
  // carry this draw keep their register budget)
  if constexpr (P <= 6) {
    if (lane != 0) return;
    double S[P][P], inv[P][P], B[P][P], y[P];
    wishart_algebra<P>(pb, chain, T, A, tagged, S, inv, B, y);
  } else {
    __shared__ double S[P * P], inv[P * P], B[P * P], Y[P * P];
    wishart_algebra_wave(pb, chain, P, T, &A[0][0], tagged, S, inv, B, Y, lane);
  }
}

// SIDE = false: after k_pairs (PAIR_BLOCKS partials). SIDE = true: after the
// colour-1 kernel of a full sweep (one partial per colour-1 block), on the
// context's second stream.
template <int P, bool SIDE>
__global__ __launch_bounds__(256) void k_wishart(DevProblem pb, uint32_t sweep)
{
  sweep += *pb.sweep_base;
  constexpr int NE = P * (P + 1) / 2;
  const int chain = blockIdx.x;
  const int nparts = SIDE ? (pb.V - pb.n0 + 255) / 256 : PAIR_BLOCKS;
  wishart_block<P>(pb, chain, nparts, pb.partials + (size_t)chain * pb.max_parts * NE, sweep);
}

// Adds the pair terms of voxel j (new coefficients bj) with neighbour k
// (coefficients bk). The reference tests b^a != 0 at the +x/+y/+z member of
// the pair (covarGPU.cu:973-988): that is k when k lies in the + direction,
// j otherwise.
template <int P>
__device__ __forceinline__ void add_pair(const float (&bj)[P], const float (&bk)[P], bool k_is_plus,
                                         double (&acc)[P * (P + 1) / 2])
{
  float df[P];
#pragma unroll
  for (int c = 0; c < P; c++) df[c] = bj[c] - bk[c];
  int e = 0;
#pragma unroll
  for (int a = 0; a < P; a++) {
    const bool ok = (k_is_plus ? bk[a] : bj[a]) != 0.0f;
    const double da = ok ? (double)df[a] : 0.0;
#pragma unroll
    for (int c = 0; c <= a; c++, e++) acc[e] = fma(da, (double)df[c], acc[e]);
  }
}

// ---------------------------------------------------------------------------
// Posterior accumulation of one saved draw per voxel (k_accum describes the
// sums). Phi is Abramowitz & Stegun 26.2.17 (|error| < 7.5e-8), as the
// reference's ProbSDNorm (covar.cpp:327-348).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double phi_as(double x)
{
  const double t = 1.0 / (1.0 + 0.2316419 * fabs(x));
  const double poly = t * (0.319381530 + t * (-0.356563782 + t * (1.781477937 + t * (-1.821255978 + t * 1.330274429))));
  const double tail = 0.39894228 * exp(-0.5 * x * x) * poly;
  return x >= 0.0 ? 1.0 - tail : tail;
}

// One voxel's save (coefficients b of sorted voxel s after this sweep's
// update): the colour updates call it for their own voxels on save sweeps,
// k_accum for every voxel.
template <int P>
__device__ __forceinline__ void accum_voxel(const DevProblem &pb, const Accum &ac, int batch, int chain, int s,
                                            const float (&b)[P])
{
  const double bw = (double)pb.bwm[chain] * (double)pb.wm[s];
  const size_t base = (size_t)chain * pb.V + s;
  const int K = P + ac.ntypes;
  // batch < 0: the draw lies past the last batch and adds to the sums only
  const bool inb = batch >= 0;
  float *bat = ac.batch + (((size_t)chain * ac.nbatches + (inb ? batch : 0)) * pb.V + s) * K;
#pragma unroll
  for (int c = 0; c < P; c++) {
    const double v = (double)b[c];
    ac.sum[base * P + c] += v;
    ac.sq[base * P + c] += v * v;
    if (inb) bat[c] += b[c];
  }
  for (int t = 0; t < ac.ntypes && t < P; t++) {
    const double pr = phi_as((double)b[t] + bw);
    ac.prb[base * ac.ntypes + t] += pr;
    if (inb) bat[P + t] += (float)pr;
  }
  ac.bwm[base] += bw;
}

template <int P>
__device__ __forceinline__ void accum_body(const DevProblem &pb, const Accum &ac, int batch, int chain, int blk)
{
  const int s = blk * 256 + threadIdx.x;
  if (s >= pb.V) return;
  const float *bp = pb.beta + ((size_t)chain * pb.V + s) * P;
  float b[P];
#pragma unroll
  for (int c = 0; c < P; c++) b[c] = bp[c];
  accum_voxel<P>(pb, ac, batch, chain, s, b);
}

// ---------------------------------------------------------------------------
// Coefficient update of one colour. With Q = X'X + n_j Lambda and
// r = m + Lambda sumV, the reference draws b = Q^-1 r + chol_lower(Q^-1) z
// (covarGPU.cu:354-392). Factor Q = U U' with U upper (the Cholesky factor of
// the index-reversed Q); then chol_lower(Q^-1) = U^-T and
// b = U^-T (U^-1 r + z): one factorisation and two triangular solves, no
// explicit inverse. PREC (colour 1 of a full sweep) also accumulates the pair
// terms of each voxel with its six colour-0 neighbours -- every neighbour pair
// has exactly one colour-1 member -- into one FP64 partial per block, which
// the CAR-precision draw (wishart_block) sums in block order.
// ---------------------------------------------------------------------------
template <int P, typename T>
__device__ __forceinline__ bool ul_draw(const T (&Q)[P][P], const T (&r)[P], const T (&z)[P],
                                        T (&out)[P], T pivot_min)
{
  T U[P][P];
  T dinv[P];
#pragma unroll
  for (int k = P - 1; k >= 0; k--) {
    T d = Q[k][k];
#pragma unroll
    for (int j = k + 1; j < P; j++) d -= U[k][j] * U[k][j];
    if (!(d >= pivot_min)) return false;
    T ukk, inv;
    if constexpr (sizeof(T) == 4) {
      // pivots are >= pivot_min (no denormals): hardware square root and
      // reciprocal, 1 ulp, in place of the library's correctly rounded forms
      ukk = __builtin_amdgcn_sqrtf(d);
      inv = __builtin_amdgcn_rcpf(ukk);
    } else {
      ukk = sqrt(d);
      inv = T(1) / ukk;
    }
    U[k][k] = ukk;
    dinv[k] = inv;
#pragma unroll
    for (int i = 0; i < k; i++) {
      T v = Q[i][k];
#pragma unroll
      for (int j = k + 1; j < P; j++) v -= U[i][j] * U[k][j];
      U[i][k] = v * dinv[k];
    }
  }
  T w[P];
#pragma unroll
  for (int i = P - 1; i >= 0; i--) {
    T v = r[i];
#pragma unroll
    for (int j = i + 1; j < P; j++) v -= U[i][j] * w[j];
    w[i] = v * dinv[i];
  }
#pragma unroll
  for (int i = 0; i < P; i++) w[i] += z[i];
#pragma unroll
  for (int i = 0; i < P; i++) {
    T v = w[i];
#pragma unroll
    for (int j = 0; j < i; j++) v -= U[j][i] * out[j];
    out[i] = v * dinv[i];
  }
  return true;
}

// FP64 refactorisation of one voxel's update (its FP32 pivot fell below
// pivot_min), P <= 6: inputs and result by value, in a function of its own,
// so that its registers do not count against the FP32 path that nearly every
// voxel takes. ok = 0: the precision is not positive definite.
template <int P>
struct CoefIn {
  float m[P], sv[P], z[P];
  float nn;
};
template <int P>
struct CoefOut {
  float out[P];
  int ok;
};
template <int P>
__device__ __forceinline__ CoefOut<P> coef_fp64(CoefIn<P> in, const float *lam, const float *xtx)
{
  double Qd[P][P], rd[P], zd[P], od[P];
#pragma unroll
  for (int i = 0; i < P; i++) {
    double acc = (double)in.m[i];
#pragma unroll
    for (int j = 0; j < P; j++) {
      double lij = (double)lam[i * P + j];
      acc += lij * (double)in.sv[j];
      Qd[i][j] = (double)xtx[i * P + j] + (double)in.nn * lij;
    }
    rd[i] = acc;
    zd[i] = (double)in.z[i];
  }
  CoefOut<P> o;
  o.ok = ul_draw<P, double>(Qd, rd, zd, od, 0.0) ? 1 : 0;
#pragma unroll
  for (int c = 0; c < P; c++) o.out[c] = (float)od[c];
  return o;
}

// FP64 refactorisation at P = 7, 8: the arithmetic of ul_draw<P, double>
// (the same operations in the same order), with every matrix and vector in an
// LDS area of the wave, one lane at a time. Such voxels are rare; keeping this
// path out of registers keeps the FP32 path (and the latent draws that share
// its kernels) within their register budget. ok = false: the precision is not
// positive definite.
template <int P>
constexpr int fp64_area() { return 2 * P * P + 6 * P; }  // Q, U, dinv, r, z, w, out

template <int P>
__device__ __noinline__ bool ul_draw_f64_lds(double *a)
{
  double *Q = a, *U = a + P * P, *dinv = U + P * P, *r = dinv + P, *z = r + P, *w = z + P, *out = w + P;
#pragma unroll 1
  for (int k = P - 1; k >= 0; k--) {
    double d = Q[k * P + k];
    for (int j = k + 1; j < P; j++) d -= U[k * P + j] * U[k * P + j];
    if (!(d >= 0.0)) return false;
    const double ukk = sqrt(d);
    const double inv = 1.0 / ukk;
    U[k * P + k] = ukk;
    dinv[k] = inv;
#pragma unroll 1
    for (int i = 0; i < k; i++) {
      double v = Q[i * P + k];
      for (int j = k + 1; j < P; j++) v -= U[i * P + j] * U[k * P + j];
      U[i * P + k] = v * dinv[k];
    }
  }
#pragma unroll 1
  for (int i = P - 1; i >= 0; i--) {
    double v = r[i];
    for (int j = i + 1; j < P; j++) v -= U[i * P + j] * w[j];
    w[i] = v * dinv[i];
  }
#pragma unroll 1
  for (int i = 0; i < P; i++) w[i] += z[i];
#pragma unroll 1
  for (int i = 0; i < P; i++) {
    double v = w[i];
    for (int j = 0; j < i; j++) v -= U[j * P + i] * out[j];
    out[i] = v * dinv[i];
  }
  return true;
}

// Every lane of the wave with ok32 = false redoes its voxel in FP64 in turn
// (wave-uniform loop over those lanes): Q = X'X + n_j Lambda and
// r = m + Lambda s_j formed in FP64 from the FP32 inputs, then
// ul_draw_f64_lds. On success out receives the FP64 draw rounded to FP32;
// otherwise out keeps `prev` (the voxel's coefficients before the update) and
// flags[1] counts the voxel. flags[0] counts every refactored voxel.
template <int P>
__device__ __forceinline__ void coef_fp64_redo(bool ok32, bool active, const float (&m)[P], const float (&sv)[P],
                                               const float (&z)[P], float nn, const float *lam, const float *xtx,
                                               const float *prev, uint32_t *flags, float (&out)[P])
{
  __shared__ double area[4][fp64_area<P>()];
  double *a = area[(threadIdx.x >> 6) & 3];
  uint64_t need = __ballot(!ok32);
  while (need) {
    const int l = __builtin_ctzll(need);
    need &= need - 1;
    if ((int)(threadIdx.x & 63) == l) {
      double *Q = a, *r = a + 2 * P * P + P, *zz = r + P, *od = r + 3 * P;
      for (int i = 0; i < P; i++) {
        double acc = (double)m[i];
        for (int j = 0; j < P; j++) {
          const double lij = (double)lam[i * P + j];
          acc += lij * (double)sv[j];
          Q[i * P + j] = (double)xtx[i * P + j] + (double)nn * lij;
        }
        r[i] = acc;
        zz[i] = (double)z[i];
      }
      if (active) atomicAdd(&flags[0], 1u);
      if (ul_draw_f64_lds<P>(a)) {
        for (int c = 0; c < P; c++) out[c] = (float)od[c];
      } else {
        // not positive definite: keep the previous coefficients
        if (active) atomicAdd(&flags[1], 1u);
        for (int c = 0; c < P; c++) out[c] = prev[c];
      }
    }
  }
}

// Block `blk` of colour `colour` (256 voxels of its list) for chain `chain`.
// KEEPNB (colour 1): the six neighbour coefficient vectors stay in registers
// for the pair terms; otherwise the pair terms read them again (the same
// values: nothing writes the other colour meanwhile), for kernels whose
// register budget is shared with the latent draws.
// The CAR precision drawn by another block of the same launch (k_overlap):
// its entries carry tag 1 beside their bits in lamtag, so the block copies
// them to LDS once every entry shows the tag. The wait is bounded (2^20
// polls, or pb.lam_spin); past the bound flags[2] is set (the host reports
// it) and the block goes on.
constexpr int LAM_SPIN_MAX = 1 << 20;
template <int P>
__device__ __forceinline__ const float *lam_wait(const DevProblem &pb, const unsigned long long *lamw,
                                                 uint32_t want = 1u)
{
  __shared__ float lam_s[P * P];
  if (threadIdx.x < P * P) {
    const int spin = pb.lam_spin > 0 ? pb.lam_spin : LAM_SPIN_MAX;
    unsigned long long v = 0ull;
    for (int it = 0;; it++) {
      v = __hip_atomic_load(lamw + threadIdx.x, __ATOMIC_RELAXED, TAG_SCOPE);
      if ((uint32_t)(v >> 32) == want) break;
      if (it >= spin && pb.pdiag && want != 1u &&
          atomicExch(pb.pdiag + 4 + gridDim.x + blockIdx.x, pb.pdiag[4 + blockIdx.x]) != pb.pdiag[4 + blockIdx.x]) {
        const uint32_t slot = atomicAdd(&pb.flags[3], 1u);
        if (slot < PT_DIAG_RECS) {
          unsigned long long *r = pb.pdiag + 4 + 2 * gridDim.x + 4 * slot;
          r[0] = (unsigned long long)(lamw + threadIdx.x);
          r[1] = want;
          r[2] = v;
          r[3] = pb.pdiag[4 + blockIdx.x];
        }
      }
      if (it >= spin || (pb.lam_spin == 0 && (it & 31) == 31 && want != 1u && tg_abort(pb))) {
        atomicOr(&pb.flags[2], 1u);
        break;
      }
      __builtin_amdgcn_s_sleep(8);
    }
    lam_s[threadIdx.x] = __uint_as_float((uint32_t)v);
  }
  __syncthreads();
  return lam_s;
}

// lamw != null: Lambda is read through lam_wait, after this block's
// neighbour and X'Z loads are under way.
// TG (the persistent sweep, one chain, relative sweep pk): what other blocks
// of the launch wrote comes from the tagged words: the other colour's b
// (colour 0 reads relative sweep pk - 1, plain pb.beta at pk = 0; colour 1
// reads pk), m_j and the pieces, Lambda of pk - 1 (plain pb.lam at pk = 0)
// and, for a voxel that keeps its coefficients, its own b of pk - 1; the new
// b goes to the tagged words and, in the launch's last sweep, to pb.beta (and
// m_j of a leftover tile to pb.M): one plain store per address and launch;
// the pair partials go to the tagged words only.
template <int P, bool PREC, bool KEEPNB = PREC, bool TG = false>
__device__ __forceinline__ void coef_body(const DevProblem &pb, int colour, uint32_t sweep, int chain, int blk,
                                          int tail, const Accum &ac, int batch,
                                          const unsigned long long *lamw = nullptr, const PTags *tg = nullptr,
                                          int pk = 0)
{
  const bool fin = !TG || pk == tg->K - 1;  // TG: plain outputs in the launch's last sweep only
  static_assert(!TG || (P <= 6 && !KEEPNB), "the persistent sweep runs P <= 6 and reads the neighbours again");
  constexpr int NE = P * (P + 1) / 2;
  const int nbk = colour ? pk : pk - 1;  // TG: relative sweep of the other colour's b
  const int q = blk * 256 + threadIdx.x;
  const int nc = colour ? pb.V - pb.n0 : pb.n0;
  const bool active = q < nc;
  const int s = (active ? q : nc - 1) + (colour ? pb.n0 : 0);
  const float *betc = pb.beta + (size_t)chain * pb.V * P;
  const float *lam = pb.lam + (size_t)chain * P * P;

  float sv[P];
#pragma unroll
  for (int c = 0; c < P; c++) sv[c] = 0.0f;
  int nn = 0;
  const int2 *nbp = reinterpret_cast<const int2 *>(pb.nbr6 + (size_t)s * 6);
  int nb[6];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    int2 v = nbp[d];
    nb[2 * d] = v.x;
    nb[2 * d + 1] = v.y;
  }
  constexpr bool KB = PREC && KEEPNB;
  float bk[KB ? 6 : 1][P];  // colour 1 keeps the neighbours for its pair terms
#pragma unroll
  for (int d = 0; d < 6; d++) {
    if (nb[d] >= 0) {
      nn++;
      const float *src = betc + (size_t)nb[d] * P;
      float v[P];
      if (TG && nbk >= 0) {
        tg_waitf<P>(tg->bt + ((size_t)(nbk & 1) * pb.V + nb[d]) * P, ptag(*tg, nbk), pb, v);
      } else {
#pragma unroll
        for (int c = 0; c < P; c++) v[c] = src[c];
      }
#pragma unroll
      for (int c = 0; c < P; c++) {
        sv[c] += v[c];
        if (KB) bk[KB ? d : 0][c] = v[c];
      }
    } else if (KB) {
#pragma unroll
      for (int c = 0; c < P; c++) bk[KB ? d : 0][c] = 0.0f;
    }
  }
  // m_j = X'(Z_j - b WM_j) = X'Z_j - b WM_j X'1, with the b of this sweep;
  // a voxel of a leftover tile of its colour (tail) has its X'Z in pieces
  float *mz = pb.M + ((size_t)chain * pb.V + s) * P;
  const float bw = pb.bwm[chain] * pb.wm[s];
  float m[P];
  bool intail;
  if constexpr (TG) intail = ((tail >> colour) & 1) && vl_tail_sum_tg<P>(pb, *tg, pk, s, m);
  else intail = ((tail >> colour) & 1) && vl_tail_sum<P>(pb, chain, s, m);
  if (intail) {
    if (active && fin) {
#pragma unroll
      for (int i = 0; i < P; i++) mz[i] = m[i];
    }
  } else if constexpr (TG) {
    tg_waitf<P>(tg->mt + (size_t)s * P, ptag(*tg, pk), pb, m);
  } else {
#pragma unroll
    for (int i = 0; i < P; i++) m[i] = mz[i];
  }
#pragma unroll
  for (int i = 0; i < P; i++) m[i] = fmaf(-bw, pb.xsum[i], m[i]);
  if constexpr (TG) {
    if (pk > 0) lam = lam_wait<P>(pb, tg->lt + (size_t)((pk - 1) & 1) * P * P, ptag(*tg, pk - 1));
  } else if (lamw) {
    lam = lam_wait<P>(pb, lamw);
  }
  float Q[P][P], r[P];
#pragma unroll
  for (int i = 0; i < P; i++) {
    float acc = m[i];
#pragma unroll
    for (int j = 0; j < P; j++) {
      float lij = lam[i * P + j];
      acc = fmaf(lij, sv[j], acc);
      Q[i][j] = fmaf((float)nn, lij, pb.xtx[i * P + j]);
    }
    r[i] = acc;
  }
  float z[P];
#pragma unroll
  for (int g = 0; g < (P + 3) / 4; g++) {
    uint32_t rr[4];
    philox10((uint32_t)s, sweep, TAG_COEF | ((uint32_t)g << 4), (uint32_t)chain, pb.key0, pb.key1, rr);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      // Box-Muller with the hardware v_log (log2), v_sqrt and v_sin/v_cos
      // (argument in revolutions): R = sqrt(-2 ln2 log2 u0), angle 2 pi u1
      const float rad = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u01f(rr[2 * h])));
      const float ang = u01f(rr[2 * h + 1]);
      if (4 * g + 2 * h < P) z[4 * g + 2 * h] = rad * __builtin_amdgcn_cosf(ang);
      if (4 * g + 2 * h + 1 < P) z[4 * g + 2 * h + 1] = rad * __builtin_amdgcn_sinf(ang);
    }
  }
  float out[P];
  float *dst = pb.beta + ((size_t)chain * pb.V + s) * P;
  const bool ok32 = ul_draw<P, float>(Q, r, z, out, pb.pivot_min);
  // FP64 refactorisation of the voxels whose FP32 pivot fell below pivot_min
  if constexpr (P <= 6) {
    if (!ok32) {
      CoefIn<P> in;
#pragma unroll
      for (int i = 0; i < P; i++) {
        in.m[i] = m[i];
        in.sv[i] = sv[i];
        in.z[i] = z[i];
      }
      in.nn = (float)nn;
      if (active) atomicAdd(&pb.flags[0], 1u);
      const CoefOut<P> o = coef_fp64<P>(in, lam, pb.xtx);
      if (o.ok) {
#pragma unroll
        for (int c = 0; c < P; c++) out[c] = o.out[c];
      } else {
        // not positive definite: keep the previous coefficients
        if (active) atomicAdd(&pb.flags[1], 1u);
        if (TG && pk > 0) {
          tg_waitf<P>(tg->bt + ((size_t)((pk - 1) & 1) * pb.V + s) * P, ptag(*tg, pk - 1), pb, out);
        } else {
#pragma unroll
          for (int c = 0; c < P; c++) out[c] = dst[c];
        }
      }
    }
  } else {
    if (__ballot(!ok32)) coef_fp64_redo<P>(ok32, active, m, sv, z, (float)nn, lam, pb.xtx, dst, pb.flags, out);
  }
  if (active) {
    if (fin) {
#pragma unroll
      for (int c = 0; c < P; c++) dst[c] = out[c];
    }
    if constexpr (TG) {
#pragma unroll
      for (int c = 0; c < P; c++)
        tg_st(tg->bt + ((size_t)(pk & 1) * pb.V + s) * P + c, ptag(*tg, pk), __float_as_uint(out[c]));
    }
    if (ac.sum) accum_voxel<P>(pb, ac, batch, chain, s, out);  // a save sweep: this voxel's final b
  }
  if (!PREC) return;

  double acc[NE];
#pragma unroll
  for (int e = 0; e < NE; e++) acc[e] = 0.0;
  if (active) {
#pragma unroll
    for (int d = 0; d < 6; d++)
      if (nb[d] >= 0) {
        if (KB) {
          add_pair<P>(out, bk[KB ? d : 0], (d & 1) != 0, acc);
        } else {
          float bn[P];
          const float *src = betc + (size_t)nb[d] * P;
          if (TG && nbk >= 0) {
            tg_waitf<P>(tg->bt + ((size_t)(nbk & 1) * pb.V + nb[d]) * P, ptag(*tg, nbk), pb, bn);
          } else {
#pragma unroll
            for (int c = 0; c < P; c++) bn[c] = src[c];
          }
          add_pair<P>(out, bn, (d & 1) != 0, acc);
        }
      }
  }
  if constexpr (TG) {
    block_sum_store_tg<NE>(acc, tg->pt + (size_t)blk * 2 * NE, ptag(*tg, pk));
    return;
  }
  double *parts = pb.partials + (size_t)chain * pb.max_parts * NE;
  block_sum_store<NE>(acc, parts + (size_t)blk * NE);
}

// Colour 1's voxels in its leftover tiles: block t of the colour-0 launch's
// extra blocks adds their pieces (vl_tail_sum) and writes m_j, beside the
// colour-0 update, so the colour-1 update reads m_j like any other voxel's.
template <int P>
__device__ __forceinline__ void tail_combine_block(const DevProblem &pb, int chain, int t)
{
  const VlRange r = vl_range(pb, 1);
  if (threadIdx.x >= 64) return;
  const int s = r.base + (r.nmain + t) * 64 + threadIdx.x;
  if (s >= r.end) return;
  float m[P];
  vl_tail_sum<P>(pb, chain, s, m);
  float *mz = pb.M + ((size_t)chain * pb.V + s) * P;
#pragma unroll
  for (int c = 0; c < P; c++) mz[c] = m[c];
}

// blocks past the colour's own (colour 0 with colour 1 in tail) run
// tail_combine_block;
// ac.sum != null: a save sweep, each updated voxel is accumulated (batch)
template <int P, bool PREC>
__global__ __launch_bounds__(256) void k_coef(DevProblem pb, int colour, uint32_t sweep, int tail, Accum ac,
                                              int batch)
{
  sweep += *pb.sweep_base;
  const int ncb = ((colour ? pb.V - pb.n0 : pb.n0) + 255) / 256;
  if ((int)blockIdx.x >= ncb) {
    tail_combine_block<P>(pb, blockIdx.y, blockIdx.x - ncb);
    return;
  }
  coef_body<P, PREC>(pb, colour, sweep, blockIdx.y, blockIdx.x, tail, ac, batch);
}

// ---------------------------------------------------------------------------
// k_overlap: the sweep of each chain (the grid's second dimension) as two
// launches, each colour's update beside the other colour's latent draws.
// Z_j and X'Z_j depend on b_j alone, and each colour's update reads only the
// other colour's coefficients, so
//   phase 0: [block 0: Lambda of sweep wsweep when wish] colour 1's latent
//            draws of `sweep` (nlat blocks), colour 0's update of `sweep`;
//   phase 1: colour 0's latent draws of sweep + 1 (nlat blocks, 0 after the
//            last sweep), colour 1's update of `sweep` with its pair
//            partials, [last block: the Bartlett factor of this sweep's
//            Lambda, drawn ahead for the next phase 0; the Lambda tags cleared];
// the kernel boundaries order everything else. The latent draws are the
// critical path, so their blocks come first and all start at once; the
// update blocks take the slots that the short leftover pieces and the first
// finished tiles free. Colour 0's update needs the Lambda that block 0 of
// the same launch draws: it reads it by its tags (lam_wait), with no fence.
// Every variate and sum is the one the serial sweep forms.
// ---------------------------------------------------------------------------
// MC: several chains (whole tiles); else one chain (colour tiles and pieces),
// compiled apart so that each keeps its own registers. At P = 7, 8 the
// colour updates beside the draws (the FP32 factor, the pair terms' 28 or 36
// FP64 sums) need more than the draws' 64-72 registers: four blocks per
// compute unit give them 128 and no spills (at C3 the X staging in LDS
// already holds a compute unit to about four blocks)
template <int P, bool MC, bool BX = false>
__global__ __launch_bounds__(256, P <= 6 ? 8 : 4) void k_overlap(DevProblem pb, int phase, uint32_t sweep, int nlat,
                                                               int wish, uint32_t wsweep, int tail, Accum ac,
                                                               int batch)
{
  constexpr int NE = P * (P + 1) / 2;
  const uint32_t base = *pb.sweep_base;
  const int chain = MC ? (int)blockIdx.y : 0;
  int b = blockIdx.x;
  if (phase == 0 && wish) {
    if (b == 0) {
      wishart_block<P>(pb, chain, (pb.V - pb.n0 + 255) / 256, pb.partials + (size_t)chain * pb.max_parts * NE,
                       wsweep + base, true, pb.wbart + (size_t)chain * P * P);
      return;
    }
    b--;
  }
  if (b < nlat) {
    // one chain: colour tiles and pieces; several: one whole tile per block
    const int c = phase ? 0 : 1;
    if (!MC) {
      vl_ctile_block<P, BX>(pb, sweep + base + (uint32_t)phase, chain, b, c ? 2 : 1);
    } else {
      const VlRange r = vl_range(pb, c);
      latent_vl_mfm_body<P, false, false, BX>(pb, sweep + base + (uint32_t)phase, chain, r.base, r.end, b, 1 << 30,
                                              0);
    }
    return;
  }
  b -= nlat;
  if (phase == 0) {
    coef_body<P, false>(pb, 0, sweep + base, chain, b, tail, ac, batch,
                        wish ? pb.lamtag + (size_t)chain * P * P : nullptr);
    return;
  }
  const int ncb = (pb.V - pb.n0 + 255) / 256;
  if (b < ncb) {
    coef_body<P, true, false>(pb, 1, sweep + base, chain, b, tail, ac, batch);
    return;
  }
  // the last block of the chain: Bartlett factor of this sweep's Lambda, tags cleared
  if (threadIdx.x < P * P) {
    pb.lamtag[(size_t)chain * P * P + threadIdx.x] = 0ull;
    pb.wbart[(size_t)chain * P * P + threadIdx.x] = bartlett_entry<P>(pb, chain, threadIdx.x, sweep + base);
  }
}

// ---------------------------------------------------------------------------
// k_persist: K full sweeps of one chain in one launch (P <= 6, the overlapped
// form's conditions). A fixed grid of resident blocks takes items from a
// queue (tg.head) until it runs out. Per relative sweep k the items are
//   L0(k), L1(k): a latent tile or piece of colour 0 / 1 (needs the b of its
//                 voxels from the update of k - 1);
//   U0(k), U1(k): the update of 256 voxels of colour 0 / 1 (needs its voxels'
//                 m_j, the other colour's b around them, Lambda of k - 1);
//                 U1 writes the pair partials, U0 a done word;
//   BT(k):        the Bartlett factor of k's Lambda (needs the draw of k - 2
//                 to have read that buffer);
//   W(k):         Lambda of k from every U1(k)'s partials, once every U0(k)
//                 is done;
// each writing tagged words that the items needing them wait for. The host
// lays the items out as a pattern of one window, repeated; entry = dk << 28 |
// type << 24 | index, and window w holds the items of relative sweep w - dk
// (items outside 0 .. K-1 are skipped). The host checks that every item
// comes after each item it waits for, so the earliest unfinished item always
// has its inputs and the queue drains whatever the residency. The same
// variates and sums as the overlapped and the serial sweeps.
// ---------------------------------------------------------------------------
template <int P>
__device__ void persist_bartlett(const DevProblem &pb, const PTags &tg, int k, uint32_t sweep)
{
  if (threadIdx.x >= P * P) return;
  // buffer k & 1 was last read by W(k - 2): its Lambda shows that
  if (k >= 2) {
    uint32_t w[1];
    tg_wait<1>(tg.lt + (size_t)(k & 1) * P * P, ptag(tg, k - 2), pb, w);
  }
  tg_std(tg.wb + ((size_t)(k & 1) * P * P + threadIdx.x) * 2, ptag(tg, k),
         bartlett_entry<P>(pb, 0, threadIdx.x, sweep));
}

template <int P>
__device__ void persist_wishart(const DevProblem &pb, const PTags &tg, int k)
{
  constexpr int NE = P * (P + 1) / 2;
  const uint32_t tag = ptag(tg, k);
  const int ncb0 = (pb.n0 + 255) / 256, ncb1 = (pb.V - pb.n0 + 255) / 256;
  // every colour-0 update of k has read Lambda of k - 1 (their done words)
  for (int b = threadIdx.x; b < ncb0; b += 256) {
    uint32_t w[1];
    tg_wait<1>(tg.u0d + b, tag, pb, w);
  }
  // the colour-1 partials, summed in sum_partials' order
  __shared__ double T[NE];
  __shared__ double A[P][P];
  double acc[NE];
#pragma unroll
  for (int e = 0; e < NE; e++) acc[e] = 0.0;
  for (int b = threadIdx.x; b < ncb1; b += 256) {
    const unsigned long long *src = tg.pt + (size_t)b * 2 * NE;
    constexpr int CH = 4;  // entries per round of loads
#pragma unroll
    for (int e0 = 0; e0 < NE; e0 += CH) {
      constexpr int W2 = 2 * CH;
      uint32_t w[W2];
      if (e0 + CH <= NE) {
        tg_wait<W2>(src + 2 * e0, tag, pb, w);
      } else {
#pragma unroll
        for (int i = 0; i < W2; i += 2) {
          if (e0 + i / 2 < NE) {
            uint32_t v[2];
            tg_wait<2>(src + 2 * e0 + i, tag, pb, v);
            w[i] = v[0];
            w[i + 1] = v[1];
          }
        }
      }
#pragma unroll
      for (int i = 0; i < CH; i++)
        if (e0 + i < NE) acc[e0 + i] += tg_hilo(w[2 * i], w[2 * i + 1]);
    }
  }
  block_sum_store<NE>(acc, T, NE);
  __syncthreads();
  const int lane = threadIdx.x;
  if (lane >= 64) return;
  for (int idx = lane; idx < P * P; idx += 64) {
    uint32_t w[2];
    tg_wait<2>(tg.wb + ((size_t)(k & 1) * P * P + idx) * 2, tag, pb, w);
    A[idx / P][idx % P] = tg_hilo(w[0], w[1]);
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  if (lane != 0) return;
  DevProblem q = pb;  // Lambda of k to buffer k & 1 with k's tag, and to pb.lam
  q.lamtag = tg.lt + (size_t)(k & 1) * P * P;
  // pb.lam_tag != 0 (set from the host only to test the bound): a tag that
  // no reader accepts
  q.lam_tag = (int)(pb.lam_tag ? (tag ^ 0x4000u) : tag);
  double S[P][P], inv[P][P], B[P][P], y[P];
  wishart_algebra<P>(q, 0, T, A, true, S, inv, B, y, k == tg.K - 1);
}

// the persistent sweep's one kernel argument
struct PArgs {
  DevProblem pb;
  PTags tg;
};

template <int P>
__global__ __launch_bounds__(256, 8) void k_persist(PArgs args)
{
  static_assert(P <= 6, "the persistent sweep runs P <= 6");
  // The argument is read in place, through the kernel-argument segment (its
  // only argument, at offset 0), by a pointer that the compiler must take as
  // new in every item: each item loads the fields it uses, and none of them
  // stays in scalar registers through another item's loops (they would crowd
  // the latent loop's registers). `args` itself is never named below, so no
  // copy of it is made.
  typedef const __attribute__((address_space(4))) PArgs KArgs;
  KArgs *ka = (KArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  const DevProblem &pb = *(const DevProblem *)&ka->pb;
  const PTags &tg = *(const PTags *)&ka->tg;
  const uint32_t base = *pb.sweep_base + tg.sweep0;
  extern __shared__ float4 xs[];
  stage_x_quads<P>(pb, xs);  // once for every latent item of the launch
  // the next queue entry is taken when an item ends, not ahead: an entry taken
  // early would wait behind this block's item while others need its outputs
  __shared__ uint32_t s_q;
  if (threadIdx.x == 0) s_q = atomicAdd(tg.head, 1u);
  __syncthreads();
  // wave-uniform (one value for the block): keeps the item's indexing scalar
  uint32_t g = __builtin_amdgcn_readfirstlane(s_q);
  const uint32_t total = (uint32_t)tg.windows * (uint32_t)tg.plen;
  while (g < total) {
    KArgs *kq = ka;
    asm volatile("" : "+s"(kq));
    const DevProblem &ip = *(const DevProblem *)&kq->pb;
    const PTags &it = *(const PTags *)&kq->tg;
    const uint32_t w = g / (uint32_t)it.plen;
    const uint32_t e = it.pattern[g - w * (uint32_t)it.plen];
    const int k = (int)w - (int)(e >> 28);
    const int idx = (int)(e & 0xFFFFFFu);
    if (k >= 0 && k < it.K) {
      const uint32_t sw = base + (uint32_t)k;
      if (threadIdx.x == 0) ip.pdiag[4 + blockIdx.x] = ((unsigned long long)(uint32_t)k << 32) | e;
      switch ((e >> 24) & 15u) {
      case PI_L0:
        vl_ctile_block_tg<P>(ip, it, k, sw, idx, 0);
        break;
      case PI_L1:
        vl_ctile_block_tg<P>(ip, it, k, sw, idx, 1);
        break;
      case PI_U0:
        coef_body<P, false, false, true>(ip, 0, sw, 0, idx, it.tail, Accum{}, 0, nullptr, &it, k);
        __syncthreads();
        if (threadIdx.x == 0) tg_st(it.u0d + idx, ptag(it, k), 1u);
        break;
      case PI_U1:
        coef_body<P, true, false, true>(ip, 1, sw, 0, idx, it.tail, Accum{}, 0, nullptr, &it, k);
        break;
      case PI_W:
        persist_wishart<P>(ip, it, k);
        break;
      default:
        persist_bartlett<P>(ip, it, k, sw);
        break;
      }
    }
    __syncthreads();  // the item's LDS is free again
    if (threadIdx.x == 0) s_q = atomicAdd(it.head, 1u);
    __syncthreads();
    g = __builtin_amdgcn_readfirstlane(s_q);
  }
}

__global__ void k_persist_init(uint32_t *head)
{
  if (threadIdx.x == 0) *head = 0u;
}

__global__ __launch_bounds__(256) void k_zero64(unsigned long long *p, size_t n)
{
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0ull;
}

// ---------------------------------------------------------------------------
// k_pairs (standalone CAR-precision update): T_ab = sum over mask voxels j and
// their +x,+y,+z neighbours k with b^a_k != 0 (covarGPU.cu:975) of
// (b^a_j - b^a_k)(b^b_j - b^b_k); float differences as the reference, exact
// FP64 products, FP64 sums, fixed block count and order.
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(256) void k_pairs(DevProblem pb)
{
  constexpr int NE = P * (P + 1) / 2;
  const int chain = blockIdx.y;
  const float *betc = pb.beta + (size_t)chain * pb.V * P;
  double acc[NE];
#pragma unroll
  for (int e = 0; e < NE; e++) acc[e] = 0.0;
  const int chunk = (pb.V + PAIR_BLOCKS - 1) / PAIR_BLOCKS;
  const int s0 = blockIdx.x * chunk;
  const int s1 = min(pb.V, s0 + chunk);
  for (int s = s0 + (int)threadIdx.x; s < s1; s += 256) {
    float bj[P];
#pragma unroll
    for (int c = 0; c < P; c++) bj[c] = betc[(size_t)s * P + c];
#pragma unroll
    for (int d = 1; d < 6; d += 2) {
      const int k = pb.nbr6[(size_t)s * 6 + d];
      if (k < 0) continue;
      float bk[P];
#pragma unroll
      for (int c = 0; c < P; c++) bk[c] = betc[(size_t)k * P + c];
      add_pair<P>(bj, bk, true, acc);
    }
  }
  block_sum_store<NE>(acc, pb.partials + ((size_t)chain * pb.max_parts + blockIdx.x) * NE);
}

// ---------------------------------------------------------------------------
// k_accum: posterior accumulation of one saved draw (mcmc.cpp:627-649), on the
// device. FP64 running sums of b and b^2 per voxel and coefficient, of the
// lesion probability Phi(b_t + b WM) for the first `ntypes` coefficients
// (compute_prbGPU, covarGPU.cu:1723-1747, 1814-1824), and of b WM; FP32 sums
// per batch of draws for batch-means Monte-Carlo errors.
// Phi is Abramowitz & Stegun 26.2.17 (|error| < 7.5e-8), as the reference's
// ProbSDNorm (covar.cpp:327-348).
template <int P>
__global__ __launch_bounds__(256) void k_accum(DevProblem pb, Accum ac, int batch)
{
  accum_body<P>(pb, ac, batch, blockIdx.y, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Plane means of the coefficients over the mask (alphaMean, covarGPU.cu:563-583):
// FP64 block partials in fixed order, summed on the host.
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(256) void k_cmeans(DevProblem pb, double *part)
{
  const int chain = blockIdx.y;
  double acc[P];
#pragma unroll
  for (int c = 0; c < P; c++) acc[c] = 0.0;
  const float *betc = pb.beta + (size_t)chain * pb.V * P;
  for (int s = blockIdx.x * 256 + threadIdx.x; s < pb.V; s += gridDim.x * 256)
#pragma unroll
    for (int c = 0; c < P; c++) acc[c] += (double)betc[(size_t)s * P + c];
  block_sum_store<P>(acc, part + ((size_t)chain * gridDim.x + blockIdx.x) * P);
}

// ---------------------------------------------------------------------------
// Per-subject log-likelihood under each subtype (compute_predictGPU with
// ProbSDNormGPU_prediction and reduce_for_Mpredict, covarGPU.cu:1783-1812,
// 1826-1957): for subject i and subtype t,
//   p = Phi(b_t,j + b WM_j + sum_{c >= T} b_c,j x_c,i)  (dProbSDNormGPU, A&S 26.2.17, FP32)
//   ll_it = sum_j y_ij log(p + 1e-35) + (1 - y_ij) log(1 - p + 1e-35)
// A block owns a tile of 256 voxels and its threads own subjects (blocks of
// 256); each thread keeps T FP64 sums over the tile. k_llreduce adds the tile
// partials in tile order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float phi_as_f(float x)
{
  const float sg = copysignf(1.0f, x), y = fabsf(x);
  const float t = 1.0f / fmaf(0.2316419f, y, 1.0f);
  const float poly = t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 1.330274429f, -1.821255978f), 1.781477937f),
                                       -0.356563782f), 0.319381530f);
  return (x >= 0.0f ? 1.0f : 0.0f) - sg * 0.39894228f * expf(-0.5f * y * y) * poly;
}

template <int P>
__global__ __launch_bounds__(256) void k_loglik(DevProblem pb, int T, double *part)
{
  const int chain = blockIdx.z;
  const int tile = blockIdx.x;
  const int i = blockIdx.y * 256 + threadIdx.x;
  const bool active = i < pb.N;
  const int ii = active ? i : 0;
  float xi[P];
#pragma unroll
  for (int c = 0; c < P; c++) xi[c] = pb.X[(size_t)c * pb.NS + ii];
  const float b = pb.bwm[chain];
  const float *betc = pb.beta + (size_t)chain * pb.V * P;
  double acc[MAX_P];
#pragma unroll
  for (int t = 0; t < MAX_P; t++) acc[t] = 0.0;
  const int s1 = min(pb.V, (tile + 1) * 256);
  for (int s = tile * 256; s < s1; s++) {
    float bs[P];
#pragma unroll
    for (int c = 0; c < P; c++) bs[c] = betc[(size_t)s * P + c];
    float rest = b * pb.wm[s];
#pragma unroll
    for (int c = 0; c < P; c++)
      if (c >= T) rest = fmaf(bs[c], xi[c], rest);
    const bool y = (pb.ybits[(size_t)s * pb.NW + (ii >> 5)] >> (ii & 31)) & 1u;
#pragma unroll
    for (int t = 0; t < P; t++) {
      if (t < T) {
        const float p = phi_as_f(bs[t] + rest);
        acc[t] += (double)(y ? logf(p + 1e-35f) : logf(1.0f - p + 1e-35f));
      }
    }
  }
  if (active) {
    double *dst = part + (((size_t)chain * gridDim.x + tile) * pb.N + i) * T;
#pragma unroll
    for (int t = 0; t < MAX_P; t++)
      if (t < T) dst[t] = acc[t];
  }
}

__global__ __launch_bounds__(256) void k_llreduce(int tiles, int NT, const double *part, double *out)
{
  const int chain = blockIdx.y;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= NT) return;
  const double *src = part + (size_t)chain * tiles * NT + k;
  double acc = 0.0;
  for (int t = 0; t < tiles; t++) acc += src[(size_t)t * NT];
  out[(size_t)chain * NT + k] = acc;
}

// ---------------------------------------------------------------------------
// Batches of independent analyses (their own masks, N and P): one launch per
// phase and covariate count, so each launch has the register allocation of
// its own P. Every block reads its work item (analysis, chain, tile or block)
// and runs the body a single analysis runs. Analyses keep their own device
// arrays; the problem table and the item lists live in device memory.
// ---------------------------------------------------------------------------
template <int P, bool BMF>
__global__ __launch_bounds__(256) void k_latent_multi(const DevProblem *pbs, const WorkItem *items, uint32_t sweep)
{
  const WorkItem it = items[blockIdx.x];
  const DevProblem pb = pbs[it.a];
  latent_vl_body<P, false, BMF>(pb, pb.lat_g, sweep + *pb.sweep_base, it.chain, it.blk, 1 << 30, 0);
}

// members drawn by inversion at G = 4: item blk is block blk of the member's
// own whole-range launch (colour tiles and pieces included), with the
// member's own X'Z form (FP32 FMA, or BF16 hi + lo on the matrix cores)
template <int P>
__global__ __launch_bounds__(256, P <= 7 ? 8 : 7) void k_latent_multi_mfm(const DevProblem *pbs,
                                                                          const WorkItem *items, uint32_t sweep)
{
  const WorkItem it = items[blockIdx.x];
  const DevProblem pb = pbs[it.a];
  const uint32_t sw = sweep + *pb.sweep_base;
  if (pb.xbx) {
    if (vl_ctiles(pb)) vl_ctile_block<P, true>(pb, sw, it.chain, it.blk, 3);
    else latent_vl_mfm_body<P, false, false, true>(pb, sw, it.chain, 0, pb.V, it.blk, 1 << 30, 0);
  } else {
    if (vl_ctiles(pb)) vl_ctile_block<P>(pb, sw, it.chain, it.blk, 3);
    else latent_vl_mfm_body<P, false>(pb, sw, it.chain, 0, pb.V, it.blk, 1 << 30, 0);
  }
}

// members in the precision form (Z stored, X'Z on the matrix cores): item
// blk is block blk of the member's own whole-range k_latent launch (Z stored,
// no X'Z), then of its k_xtz_mfma launch
template <int P>
__global__ __launch_bounds__(256) void k_latent_multi_z(const DevProblem *pbs, const WorkItem *items, uint32_t sweep)
{
  const WorkItem it = items[blockIdx.x];
  const DevProblem pb = pbs[it.a];
  int lshift, reps;
  lane_grouping(pb, lshift, reps);
  const uint32_t sw = sweep + *pb.sweep_base;
  if (reps == 1) latent_body<P, true, false, 0>(pb, lshift, reps, LATENT_VITER, sw, it.chain, it.blk);
  else latent_body<P, false, false, 0>(pb, lshift, reps, LATENT_VITER, sw, it.chain, it.blk);
}

template <int P>
__global__ __launch_bounds__(256) void k_xtz_mfma_multi(const DevProblem *pbs, const WorkItem *items)
{
  const WorkItem it = items[blockIdx.x];
  xtz_mfma_body<P>(pbs[it.a], it.chain, it.blk);
}

template <int P, bool PREC>
__global__ __launch_bounds__(256) void k_coef_multi(const DevProblem *pbs, const WorkItem *items, int colour,
                                                    uint32_t sweep, const Accum *accs, BatchIndex bidx)
{
  const WorkItem it = items[blockIdx.x];
  const DevProblem pb = pbs[it.a];
  // the batch's latent launch always precedes: with colour tiles, colour 0
  // adds its own leftover pieces and, in the items past its blocks, colour 1's
  // (tail_combine_block)
  const int ncb = ((colour ? pb.V - pb.n0 : pb.n0) + 255) / 256;
  if (it.blk >= ncb) {
    tail_combine_block<P>(pb, it.chain, it.blk - ncb);
    return;
  }
  const Accum ac = accs ? accs[it.a] : Accum{};
  coef_body<P, PREC>(pb, colour, sweep + *pb.sweep_base, it.chain, it.blk, (colour == 0 && vl_ctiles(pb)) ? 1 : 0,
                     ac, accs ? bidx.b[it.a] : 0);
}

template <int P>
__global__ __launch_bounds__(256) void k_wishart_multi(const DevProblem *pbs, const WorkItem *items, uint32_t sweep)
{
  constexpr int NE = P * (P + 1) / 2;
  const WorkItem it = items[blockIdx.x];
  const DevProblem pb = pbs[it.a];
  wishart_block<P>(pb, it.chain, (pb.V - pb.n0 + 255) / 256, pb.partials + (size_t)it.chain * pb.max_parts * NE,
                   sweep + *pb.sweep_base);
}


// ---------------------------------------------------------------------------
// Batches in one launch per phase: every block reads its item and runs the
// body of its member's covariate count (a switch on P), so that a batch
// sweep is four launches on one stream whatever the covariate counts. Each
// kernel's register allocation is that of its largest body (P = 8); batches
// of small analyses, which cannot fill the chip and would otherwise wait on
// one launch per phase and covariate count, gain from the fewer launches.
// ---------------------------------------------------------------------------
#define BSG_SWITCH_P(PV, BODY) \
  switch (PV) {                \
    case 1: BODY(1); break;    \
    case 2: BODY(2); break;    \
    case 3: BODY(3); break;    \
    case 4: BODY(4); break;    \
    case 5: BODY(5); break;    \
    case 6: BODY(6); break;    \
    case 7: BODY(7); break;    \
    case 8: BODY(8); break;    \
    default: break;            \
  }

template <bool BMF>
__global__ __launch_bounds__(256) void k_latent_multi_all(const DevProblem *pbs, const WorkItem *items,
                                                             uint32_t sweep)
{
  const WorkItem it = items[blockIdx.x];
  const DevProblem pb = pbs[it.a];
  const uint32_t sw = sweep + *pb.sweep_base;
  // members at four lanes per voxel drawn by inversion: the G = 4 body (colour tiles and pieces)
  const bool g4 = !BMF && pb.lat_g == 4;
#define B_(PP)                                                                                        \
  do {                                                                                                \
    if (g4 && pb.xbx) {                                                                               \
      if (vl_ctiles(pb)) vl_ctile_block<PP, true>(pb, sw, it.chain, it.blk, 3);                       \
      else latent_vl_mfm_body<PP, false, false, true>(pb, sw, it.chain, 0, pb.V, it.blk, 1 << 30, 0); \
    } else if (g4) {                                                                                  \
      if (vl_ctiles(pb)) vl_ctile_block<PP>(pb, sw, it.chain, it.blk, 3);                             \
      else latent_vl_mfm_body<PP, false>(pb, sw, it.chain, 0, pb.V, it.blk, 1 << 30, 0);              \
    } else {                                                                                          \
      latent_vl_body<PP, false, BMF>(pb, pb.lat_g, sw, it.chain, it.blk, 1 << 30, 0);                 \
    }                                                                                                 \
  } while (0)
  BSG_SWITCH_P(pb.P, B_)
#undef B_
}

template <bool PREC>
__global__ __launch_bounds__(256) void k_coef_multi_all(const DevProblem *pbs, const WorkItem *items, int colour,
                                                        uint32_t sweep, const Accum *accs, BatchIndex bidx)
{
  const WorkItem it = items[blockIdx.x];
  const DevProblem pb = pbs[it.a];
  const int ncb = ((colour ? pb.V - pb.n0 : pb.n0) + 255) / 256;
  const Accum ac = accs ? accs[it.a] : Accum{};
#define B_(PP)                                                                                             \
  do {                                                                                                     \
    if (it.blk >= ncb) tail_combine_block<PP>(pb, it.chain, it.blk - ncb);                                 \
    else coef_body<PP, PREC>(pb, colour, sweep + *pb.sweep_base, it.chain, it.blk,                        \
                             (colour == 0 && vl_ctiles(pb)) ? 1 : 0, ac, accs ? bidx.b[it.a] : 0);         \
  } while (0)
  BSG_SWITCH_P(pb.P, B_)
#undef B_
}

__global__ __launch_bounds__(256) void k_wishart_multi_all(const DevProblem *pbs, const WorkItem *items,
                                                           uint32_t sweep)
{
  const WorkItem it = items[blockIdx.x];
  const DevProblem pb = pbs[it.a];
#define B_(PP)                                                                                           \
  wishart_block<PP>(pb, it.chain, (pb.V - pb.n0 + 255) / 256,                                            \
                    pb.partials + (size_t)it.chain * pb.max_parts * ((PP) * ((PP) + 1) / 2),             \
                    sweep + *pb.sweep_base)
  BSG_SWITCH_P(pb.P, B_)
#undef B_
}

hipError_t launch_batch_all(int phase, const DevProblem *pbs, const WorkItem *items, int n, size_t lds, bool bmf,
                            uint32_t sweep, hipStream_t st, const Accum *accs, const BatchIndex &bidx)
{
  if (n <= 0) return hipSuccess;
  if (lds > 65536) return hipErrorInvalidValue;
  const dim3 grid((unsigned)n);
  if (phase == 0) {
    if (bmf) hipLaunchKernelGGL((k_latent_multi_all<true>), grid, dim3(256), lds, st, pbs, items, sweep);
    else hipLaunchKernelGGL((k_latent_multi_all<false>), grid, dim3(256), lds, st, pbs, items, sweep);
  } else if (phase == 1) {
    hipLaunchKernelGGL((k_coef_multi_all<false>), grid, dim3(256), 0, st, pbs, items, 0, sweep, accs, bidx);
  } else if (phase == 2) {
    hipLaunchKernelGGL((k_coef_multi_all<true>), grid, dim3(256), 0, st, pbs, items, 1, sweep, accs, bidx);
  } else {
    hipLaunchKernelGGL(k_wishart_multi_all, grid, dim3(256), 0, st, pbs, items, sweep);
  }
  return hipGetLastError();
}

// ===========================================================================
// Wide problems: MAX_P < P <= PW_MAX (20) covariates. The reference's GPU path
// accepts P <= 20 (per-thread var[20*20], covarGPU.cu:65, 300-304). These
// kernels take P at run time: per-voxel vectors sit in registers through
// loops unrolled to PW_MAX under a uniform predicate (c < P), a voxel's P x P
// factor sits in LDS. They carry the Z-free fused form in the serial order:
// latent draws with X'Z (and the WM statistic), colour updates (with the
// posterior accumulation on save sweeps), FP64 pair sums, the Wishart draw,
// and the per-save summaries. The random counters are those of the other
// kernels, so the draws are the same variates.
// ===========================================================================

// Latent draws + X'Z (+ the WM statistic with WMU), Z not kept. G lanes per
// voxel split its subject quads into G contiguous ranges (as latent_vl_body);
// X is read as [NQ][P] quads, staged in LDS when it fits (xlds), else from
// global memory (every lane of a subject group reads the same address).
template <bool WMU>
__global__ __launch_bounds__(256) void k_latent_w(DevProblem pb, int G, uint32_t sweep, int xlds)
{
  sweep += *pb.sweep_base;
  extern __shared__ float4 xsm[];
  const int P = pb.P;
  const float4 *xq = pb.Xq;
  if (xlds) {
    for (int i = threadIdx.x; i < pb.NQ * P; i += blockDim.x) xsm[i] = pb.Xq[i];
    __syncthreads();
    xq = xsm;
  }
  const int chain = blockIdx.y;
  const int lgg = G >= 8 ? 3 : G >= 4 ? 2 : G >= 2 ? 1 : 0;
  const int VPW = 64 >> lgg;
  const int lane = threadIdx.x & 63, vl = lane & (VPW - 1), g = lane >> (6 - lgg);
  const int R = (pb.NQ + (1 << lgg) - 1) >> lgg;
  const int q0 = g * R, q1 = min(q0 + R, pb.NQ);
  const float b = pb.bwm[chain];
  const float eps = pb.trunc_eps;
  const int tiles = (pb.V + 4 * VPW - 1) / (4 * VPW);
  double bpart = 0.0;
  for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
    const int s = (t * 4 + (threadIdx.x >> 6)) * VPW + vl;
    const bool valid = s < pb.V;
    const int sv = valid ? s : pb.V - 1;
    const float wmv = pb.wm[sv];
    const float bw = b * wmv;
    const float *bp = pb.beta + ((size_t)chain * pb.V + sv) * P;
    float bs[PW_MAX], acc[PW_MAX];
#pragma unroll
    for (int c = 0; c < PW_MAX; c++) {
      bs[c] = c < P ? bp[c] : 0.0f;
      acc[c] = 0.0f;
    }
    const uint32_t *yrow = pb.ybits + (size_t)sv * pb.NW;
    const uint32_t cbase = (uint32_t)sv * (uint32_t)pb.NQ;
    float zs = 0.0f;
    for (int q = q0; q < q1; q++) {
      const float4 *xr = xq + (size_t)q * P;
      float mu[4] = {bw, bw, bw, bw};
#pragma unroll
      for (int c = 0; c < PW_MAX; c++)
        if (c < P) {
          const float4 x = xr[c];
          mu[0] = fmaf(bs[c], x.x, mu[0]);
          mu[1] = fmaf(bs[c], x.y, mu[1]);
          mu[2] = fmaf(bs[c], x.z, mu[2]);
          mu[3] = fmaf(bs[c], x.w, mu[3]);
        }
      const uint32_t yw = yrow[q >> 3] >> ((q & 7) * 4);
      uint32_t rr[4];
      philox10(cbase + (uint32_t)q, sweep, TAG_LATENT, (uint32_t)chain, pb.key0, pb.key1, rr);
      float z[4];
      latent_quad(mu, yw, rr, eps, z);  // X is zero past N: padding subjects add nothing to m_j
      if (WMU && 4 * q + 3 >= pb.N) {
#pragma unroll
        for (int l = 0; l < 4; l++) z[l] = (4 * q + l < pb.N) ? z[l] : 0.0f;
      }
#pragma unroll
      for (int c = 0; c < PW_MAX; c++)
        if (c < P) {
          const float4 x = xr[c];
          acc[c] = fmaf(x.x, z[0], acc[c]);
          acc[c] = fmaf(x.y, z[1], acc[c]);
          acc[c] = fmaf(x.z, z[2], acc[c]);
          acc[c] = fmaf(x.w, z[3], acc[c]);
        }
      if (WMU) zs += (z[0] + z[1]) + (z[2] + z[3]);
    }
    for (int off = VPW; off < 64; off <<= 1) {
#pragma unroll
      for (int c = 0; c < PW_MAX; c++)
        if (c < P) acc[c] += __shfl_xor(acc[c], off, 64);
      if (WMU) zs += __shfl_xor(zs, off, 64);
    }
    if (valid && g == 0) {
      float *m = pb.M + ((size_t)chain * pb.V + s) * P;
#pragma unroll
      for (int c = 0; c < PW_MAX; c++)
        if (c < P) m[c] = acc[c];
      if (WMU) {
        double r = (double)zs;
        for (int c = 0; c < P; c++) r -= (double)pb.xsum[c] * (double)bs[c];
        bpart += (double)wmv * r;
      }
    }
  }
  if (WMU) {
    double bp1[1] = {bpart};
    block_sum_store<1>(bp1, pb.bstat + (size_t)chain * BSTAT_SLOTS + blockIdx.x);
  }
}

// Packed upper triangle, row-major: entry (i, j), i <= j, of a P x P matrix.
__device__ __forceinline__ int upk(int i, int j, int P) { return i * P - (i * (i - 1)) / 2 + (j - i); }

// Floats of one thread's LDS area in k_coef_w: the packed factor and a
// vector, padded to an odd stride so that the threads' areas fall in
// different banks.
__host__ __device__ inline int coefw_stride(int P) { return (P * (P + 1) / 2 + P) | 1; }
constexpr int COEFW_THREADS = 64;

// Posterior accumulation of one voxel (accum_voxel with P at run time).
__device__ __forceinline__ void accum_voxel_rt(const DevProblem &pb, const Accum &ac, int batch, int chain, int s,
                                               const float *b)
{
  const int P = pb.P;
  const double bw = (double)pb.bwm[chain] * (double)pb.wm[s];
  const size_t base = (size_t)chain * pb.V + s;
  const int K = P + ac.ntypes;
  const bool inb = batch >= 0;
  float *bat = ac.batch + (((size_t)chain * ac.nbatches + (inb ? batch : 0)) * pb.V + s) * K;
  for (int c = 0; c < P; c++) {
    const double v = (double)b[c];
    ac.sum[base * P + c] += v;
    ac.sq[base * P + c] += v * v;
    if (inb) bat[c] += b[c];
  }
  for (int t = 0; t < ac.ntypes && t < P; t++) {
    const double pr = phi_as((double)b[t] + bw);
    ac.prb[base * ac.ntypes + t] += pr;
    if (inb) bat[P + t] += (float)pr;
  }
  ac.bwm[base] += bw;
}

// The four normals of group g of voxel s's coefficient draw (coef_body's
// Box-Muller on the hardware v_log / v_sqrt / v_sin / v_cos)
__device__ __forceinline__ void coef_normals(const DevProblem &pb, int s, uint32_t sweep, int chain, int g,
                                             float (&zz)[4])
{
  uint32_t rr[4];
  philox10((uint32_t)s, sweep, TAG_COEF | ((uint32_t)g << 4), (uint32_t)chain, pb.key0, pb.key1, rr);
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const float rad = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u01f(rr[2 * h])));
    const float ang = u01f(rr[2 * h + 1]);
    zz[2 * h] = rad * __builtin_amdgcn_cosf(ang);
    zz[2 * h + 1] = rad * __builtin_amdgcn_sinf(ang);
  }
}

// The coefficient update of one voxel in FP64, for one lane at a time (the
// wave-uniform loop of k_coef_w): its FP32 inputs formed again from memory as
// k_coef_w forms them (the neighbour sum s_j and m_j - b WM_j X'1), then the
// factor and solves of k_coef_w's FP32 path in FP64 in the wave's LDS area
// (packed factor, then a vector, which holds the result). Returns false if
// the precision is not positive definite.
__device__ bool coef_w_fp64(const DevProblem &pb, int s, uint32_t sweep, int chain, double *U)
{
  const int P = pb.P, NU = P * (P + 1) / 2;
  double *t = U + NU;
  const float *betc = pb.beta + (size_t)chain * pb.V * P;
  const float *lam = pb.lam + (size_t)chain * P * P;
  const float bw = pb.bwm[chain] * pb.wm[s];
  const float *mz = pb.M + ((size_t)chain * pb.V + s) * P;
  int nn = 0;
  for (int d = 0; d < 6; d++) nn += pb.nbr6[(size_t)s * 6 + d] >= 0 ? 1 : 0;
  // U[j] (j < P) first holds s_j in FP32, summed in neighbour order
  float *svf = reinterpret_cast<float *>(U + NU + P);
  for (int c = 0; c < P; c++) svf[c] = 0.0f;
  for (int d = 0; d < 6; d++) {
    const int k = pb.nbr6[(size_t)s * 6 + d];
    if (k >= 0)
      for (int c = 0; c < P; c++) svf[c] += betc[(size_t)k * P + c];
  }
  for (int i = 0; i < P; i++) {
    double acc = (double)fmaf(-bw, pb.xsum[i], mz[i]);
    for (int j = 0; j < P; j++) acc += (double)lam[i * P + j] * (double)svf[j];
    t[i] = acc;
  }
  for (int i = 0; i < P; i++)
    for (int j = i; j < P; j++) U[upk(i, j, P)] = (double)pb.xtx[i * P + j] + (double)nn * (double)lam[i * P + j];
  for (int k = P - 1; k >= 0; k--) {
    double d = U[upk(k, k, P)];
    for (int j = k + 1; j < P; j++) d -= U[upk(k, j, P)] * U[upk(k, j, P)];
    if (!(d >= 0.0)) return false;
    const double ukk = sqrt(d);
    U[upk(k, k, P)] = ukk;
    for (int i = 0; i < k; i++) {
      double v = U[upk(i, k, P)];
      for (int j = k + 1; j < P; j++) v -= U[upk(i, j, P)] * U[upk(k, j, P)];
      U[upk(i, k, P)] = v / ukk;
    }
  }
  for (int i = P - 1; i >= 0; i--) {
    double v = t[i];
    for (int j = i + 1; j < P; j++) v -= U[upk(i, j, P)] * t[j];
    t[i] = v / U[upk(i, i, P)];
  }
  for (int g = 0; 4 * g < P; g++) {
    float zz[4];
    coef_normals(pb, s, sweep, chain, g, zz);
    for (int h = 0; h < 4 && 4 * g + h < P; h++) t[4 * g + h] += (double)zz[h];
  }
  for (int i = 0; i < P; i++) {
    double v = t[i];
    for (int j = 0; j < i; j++) v -= U[upk(j, i, P)] * t[j];
    t[i] = v / U[upk(i, i, P)];
  }
  return true;
}

// Coefficient update of one colour, one voxel per thread, 64 threads (one
// wave) per block. The neighbour sum s_j, then Q = X'X + n_j Lambda (upper
// triangle) factored in place as Q = U U' with U upper (ul_draw's scheme),
// r = m + Lambda s_j solved through U and U' with the normals added between:
// b = U^-T (U^-1 r + z). A pivot below pivot_min sends the voxel to the FP64
// form (coef_w_fp64); if that fails too, the voxel keeps its coefficients
// and flags[1] counts it.
__global__ __launch_bounds__(64) void k_coef_w(DevProblem pb, int colour, uint32_t sweep, Accum ac, int batch)
{
  sweep += *pb.sweep_base;
  const int P = pb.P, NU = P * (P + 1) / 2;
  extern __shared__ float wsm[];
  float *U = wsm + threadIdx.x * coefw_stride(P);
  float *t = U + NU;
  // the FP64 redo's area: packed factor, vector, and the FP32 neighbour sum
  __shared__ double f64area[PW_MAX * (PW_MAX + 1) / 2 + PW_MAX + PW_MAX / 2];
  const int chain = blockIdx.y;
  const int q = blockIdx.x * COEFW_THREADS + threadIdx.x;
  const int nc = colour ? pb.V - pb.n0 : pb.n0;
  const bool active = q < nc;
  const int s = (active ? q : nc - 1) + (colour ? pb.n0 : 0);
  const float *betc = pb.beta + (size_t)chain * pb.V * P;
  const float *lam = pb.lam + (size_t)chain * P * P;
  int nb[6];
#pragma unroll
  for (int d = 0; d < 6; d++) nb[d] = pb.nbr6[(size_t)s * 6 + d];
  float sv[PW_MAX], m[PW_MAX];
  int nn = 0;
#pragma unroll
  for (int c = 0; c < PW_MAX; c++) sv[c] = 0.0f;
#pragma unroll
  for (int d = 0; d < 6; d++)
    if (nb[d] >= 0) {
      nn++;
      const float *src = betc + (size_t)nb[d] * P;
#pragma unroll
      for (int c = 0; c < PW_MAX; c++)
        if (c < P) sv[c] += src[c];
    }
  const float bw = pb.bwm[chain] * pb.wm[s];
  const float *mz = pb.M + ((size_t)chain * pb.V + s) * P;
#pragma unroll
  for (int c = 0; c < PW_MAX; c++) m[c] = c < P ? fmaf(-bw, pb.xsum[c], mz[c]) : 0.0f;
  for (int i = 0; i < P; i++) {
    float acc = m[0];
#pragma unroll
    for (int c = 0; c < PW_MAX; c++) acc = (c == i) ? m[c] : acc;
#pragma unroll
    for (int j = 0; j < PW_MAX; j++)
      if (j < P) acc = fmaf(lam[i * P + j], sv[j], acc);
    t[i] = acc;
  }
  for (int i = 0; i < P; i++)
    for (int j = i; j < P; j++) U[upk(i, j, P)] = fmaf((float)nn, lam[i * P + j], pb.xtx[i * P + j]);
  bool ok = true;
  for (int k = P - 1; k >= 0; k--) {
    float d = U[upk(k, k, P)];
    for (int j = k + 1; j < P; j++) {
      const float u = U[upk(k, j, P)];
      d -= u * u;
    }
    if (!(d >= pb.pivot_min)) {
      ok = false;
      break;
    }
    // pivots are >= pivot_min (no denormals): hardware square root and reciprocal
    const float ukk = __builtin_amdgcn_sqrtf(d);
    const float inv = __builtin_amdgcn_rcpf(ukk);
    U[upk(k, k, P)] = ukk;
    for (int i = 0; i < k; i++) {
      float v = U[upk(i, k, P)];
      for (int j = k + 1; j < P; j++) v -= U[upk(i, j, P)] * U[upk(k, j, P)];
      U[upk(i, k, P)] = v * inv;
    }
  }
  float out[PW_MAX];
  if (ok) {
    for (int i = P - 1; i >= 0; i--) {
      float v = t[i];
      for (int j = i + 1; j < P; j++) v -= U[upk(i, j, P)] * t[j];
      t[i] = v * __builtin_amdgcn_rcpf(U[upk(i, i, P)]);
    }
    for (int g = 0; 4 * g < P; g++) {
      float zz[4];
      coef_normals(pb, s, sweep, chain, g, zz);
      for (int h = 0; h < 4 && 4 * g + h < P; h++) t[4 * g + h] += zz[h];
    }
    for (int i = 0; i < P; i++) {
      float v = t[i];
      for (int j = 0; j < i; j++) v -= U[upk(j, i, P)] * t[j];
      t[i] = v * __builtin_amdgcn_rcpf(U[upk(i, i, P)]);
    }
#pragma unroll
    for (int c = 0; c < PW_MAX; c++) out[c] = c < P ? t[c] : 0.0f;
  }
  float *dst = pb.beta + ((size_t)chain * pb.V + s) * P;
  // FP64 refactorisation, one lane at a time in the wave's area
  uint64_t need = __ballot(!ok);
  while (need) {
    const int l = __builtin_ctzll(need);
    need &= need - 1;
    if ((int)threadIdx.x == l) {
      if (active) atomicAdd(&pb.flags[0], 1u);
      const bool ok64 = coef_w_fp64(pb, s, sweep, chain, f64area);
      if (!ok64 && active) atomicAdd(&pb.flags[1], 1u);  // not positive definite: keep the coefficients
      const double *t64 = f64area + NU;
#pragma unroll
      for (int c = 0; c < PW_MAX; c++)
        if (c < P) out[c] = ok64 ? (float)t64[c] : dst[c];
    }
  }
  if (active) {
#pragma unroll
    for (int c = 0; c < PW_MAX; c++)
      if (c < P) dst[c] = out[c];
    if (ac.sum) accum_voxel_rt(pb, ac, batch, chain, s, dst);  // a save sweep: this voxel's final b
  }
}

// FP64 pair sums (k_pairs with P at run time): PAIR_BLOCKS blocks per chain,
// each over a contiguous range of voxels in rounds of 64. A round stages the
// float differences of each voxel's +x, +y, +z pairs in LDS (the a-side with
// the reference's b^a_k != 0 test applied, covarGPU.cu:975); thread e owns
// entry e = (a, c), c <= a, and adds the exact FP64 products of the round in
// pair order. One partial per block and entry.
constexpr int PAIRW_ROWS = 192;
__global__ __launch_bounds__(256) void k_pairs_w(DevProblem pb)
{
  const int P = pb.P, NE = P * (P + 1) / 2;
  __shared__ float dA[PAIRW_ROWS][PW_MAX], dC[PAIRW_ROWS][PW_MAX];
  const int chain = blockIdx.y;
  const float *betc = pb.beta + (size_t)chain * pb.V * P;
  const int chunk = (pb.V + PAIR_BLOCKS - 1) / PAIR_BLOCKS;
  const int s0 = blockIdx.x * chunk;
  const int s1 = min(pb.V, s0 + chunk);
  const int e = threadIdx.x;
  int ea = 0, ec = 0;  // e = ea (ea + 1) / 2 + ec, ec <= ea
  while ((ea + 1) * (ea + 2) / 2 <= e) ea++;
  ec = e - ea * (ea + 1) / 2;
  double acc = 0.0;
  for (int base = s0; base < s1; base += PAIRW_ROWS / 3) {
    if (threadIdx.x < PAIRW_ROWS) {
      const int row = threadIdx.x, sidx = base + row / 3, d = row % 3;
      const int k = sidx < s1 ? pb.nbr6[(size_t)sidx * 6 + 2 * d + 1] : -1;
      for (int c = 0; c < P; c++) {
        float da = 0.0f, dc = 0.0f;
        if (k >= 0) {
          const float bk = betc[(size_t)k * P + c];
          dc = betc[(size_t)sidx * P + c] - bk;
          da = bk != 0.0f ? dc : 0.0f;
        }
        dA[row][c] = da;
        dC[row][c] = dc;
      }
    }
    __syncthreads();
    if (e < NE)
      for (int row = 0; row < PAIRW_ROWS; row++) acc = fma((double)dA[row][ea], (double)dC[row][ec], acc);
    __syncthreads();
  }
  if (e < NE) pb.partials[((size_t)chain * pb.max_parts + blockIdx.x) * NE + e] = acc;
}

// Bartlett entry idx = i P + j with P at run time (bartlett_entry)
__device__ double bartlett_entry_rt(const DevProblem &pb, int chain, int idx, uint32_t sweep)
{
  const int P = pb.P, i = idx / P, j = idx % P;
  double val = 0.0;
  if (i == j) {
    val = sqrt(2.0 * gamma_mt(0.5 * (double)(pb.df - i), (uint32_t)i, sweep, (uint32_t)chain, pb.key0, pb.key1));
  } else if (j < i) {
    uint32_t rr[4];
    philox10((uint32_t)idx, sweep, TAG_WNORM, (uint32_t)chain, pb.key0, pb.key1, rr);
    val = sqrt(-2.0 * log(u01d(rr[0]))) * cospi(2.0 * u01d(rr[1]));
  }
  return val;
}

// The CAR-precision draw with P at run time (wishart_block's steps): thread e
// sums entry e of the PAIR_BLOCKS partials in block order, the Bartlett factor
// is drawn across the block, then the first wave forms S = (T + I)^-1,
// L = chol(S) and Lambda = (L A)(L A)' in FP64 with the matrices in LDS.
__global__ __launch_bounds__(256) void k_wishart_w(DevProblem pb, uint32_t sweep)
{
  sweep += *pb.sweep_base;
  const int P = pb.P, NE = P * (P + 1) / 2;
  const int chain = blockIdx.x;
  __shared__ double T[PW_MAX * (PW_MAX + 1) / 2];
  __shared__ double A[PW_MAX * PW_MAX], S[PW_MAX * PW_MAX], inv[PW_MAX * PW_MAX], B[PW_MAX * PW_MAX],
      Y[PW_MAX * PW_MAX];
  const double *part = pb.partials + (size_t)chain * pb.max_parts * NE;
  if ((int)threadIdx.x < NE) {
    double acc = 0.0;
    for (int b = 0; b < PAIR_BLOCKS; b++) acc += part[(size_t)b * NE + threadIdx.x];
    T[threadIdx.x] = acc;
  }
  for (int idx = threadIdx.x; idx < P * P; idx += blockDim.x) A[idx] = bartlett_entry_rt(pb, chain, idx, sweep);
  __syncthreads();
  if (threadIdx.x >= 64) return;
  wishart_algebra_wave(pb, chain, P, T, A, false, S, inv, B, Y, threadIdx.x);
}

__global__ __launch_bounds__(256) void k_accum_w(DevProblem pb, Accum ac, int batch)
{
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= pb.V) return;
  const int chain = blockIdx.y;
  accum_voxel_rt(pb, ac, batch, chain, s, pb.beta + ((size_t)chain * pb.V + s) * pb.P);
}

__global__ __launch_bounds__(256) void k_cmeans_w(DevProblem pb, double *part)
{
  const int chain = blockIdx.y, P = pb.P;
  double acc[PW_MAX];
#pragma unroll
  for (int c = 0; c < PW_MAX; c++) acc[c] = 0.0;
  const float *betc = pb.beta + (size_t)chain * pb.V * P;
  for (int s = blockIdx.x * 256 + threadIdx.x; s < pb.V; s += gridDim.x * 256)
#pragma unroll
    for (int c = 0; c < PW_MAX; c++)
      if (c < P) acc[c] += (double)betc[(size_t)s * P + c];
  block_sum_store<PW_MAX>(acc, part + ((size_t)chain * gridDim.x + blockIdx.x) * P, P);
}

// per-subject log-likelihood (k_loglik with P at run time)
__global__ __launch_bounds__(256) void k_loglik_w(DevProblem pb, int T, double *part)
{
  const int P = pb.P;
  const int chain = blockIdx.z;
  const int tile = blockIdx.x;
  const int i = blockIdx.y * 256 + threadIdx.x;
  const bool active = i < pb.N;
  const int ii = active ? i : 0;
  float xi[PW_MAX];
#pragma unroll
  for (int c = 0; c < PW_MAX; c++) xi[c] = c < P ? pb.X[(size_t)c * pb.NS + ii] : 0.0f;
  const float b = pb.bwm[chain];
  const float *betc = pb.beta + (size_t)chain * pb.V * P;
  double acc[PW_MAX];
#pragma unroll
  for (int t = 0; t < PW_MAX; t++) acc[t] = 0.0;
  const int s1 = min(pb.V, (tile + 1) * 256);
  for (int s = tile * 256; s < s1; s++) {
    float bs[PW_MAX];
#pragma unroll
    for (int c = 0; c < PW_MAX; c++) bs[c] = c < P ? betc[(size_t)s * P + c] : 0.0f;
    float rest = b * pb.wm[s];
#pragma unroll
    for (int c = 0; c < PW_MAX; c++)
      if (c >= T && c < P) rest = fmaf(bs[c], xi[c], rest);
    const bool y = (pb.ybits[(size_t)s * pb.NW + (ii >> 5)] >> (ii & 31)) & 1u;
#pragma unroll
    for (int t = 0; t < PW_MAX; t++) {
      if (t < T) {
        const float p = phi_as_f(bs[t] + rest);
        acc[t] += (double)(y ? logf(p + 1e-35f) : logf(1.0f - p + 1e-35f));
      }
    }
  }
  if (active) {
    double *dst = part + (((size_t)chain * gridDim.x + tile) * pb.N + i) * T;
#pragma unroll
    for (int t = 0; t < PW_MAX; t++)
      if (t < T) dst[t] = acc[t];
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------


#define BSG_DISPATCH_P(PVAL, CALL) \
  switch (PVAL) {                  \
    case 1: CALL(1); break;        \
    case 2: CALL(2); break;        \
    case 3: CALL(3); break;        \
    case 4: CALL(4); break;        \
    case 5: CALL(5); break;        \
    case 6: CALL(6); break;        \
    case 7: CALL(7); break;        \
    case 8: CALL(8); break;        \
    default: return hipErrorInvalidValue; \
  }

// lanes per voxel of the wide latent kernel: a set split, else enough lanes
// per voxel to give the chip many waves, few enough that every lane keeps
// several quads
static int wide_groups(const DevProblem &pb)
{
  if (pb.lat_g > 0) return pb.lat_g;
  if (pb.NQ <= 4) return 1;
  if (pb.NQ <= 16) return 2;
  if (pb.NQ <= 128) return 4;
  return 8;
}

// the lane split chosen from N (vl_groups without a set split)
int latent_vl_groups(const DevProblem &pb)
{
  DevProblem q = pb;
  q.lat_g = 0;
  return pb.P > MAX_P ? wide_groups(q) : vl_groups(q);
}

// ---- launchers of the wide kernels (MAX_P < P <= PW_MAX) ----
static hipError_t launch_latent_w(const DevProblem &pb, uint32_t sweep, hipStream_t st, int *parts)
{
  const int G = wide_groups(pb);
  if (G != 1 && G != 2 && G != 4 && G != 8) return hipErrorInvalidValue;
  const int VPW = 64 / G;
  const long long tiles = ((long long)pb.V + 4 * VPW - 1) / (4 * VPW);
  const int nlb = (int)std::min<long long>(tiles, BSTAT_SLOTS);
  const size_t xbytes = (size_t)pb.NQ * pb.P * sizeof(float4);
  const int xlds = xbytes <= 65536 ? 1 : 0;
  dim3 grid((unsigned)nlb, (unsigned)pb.nchains);
  if (pb.wm_update)
    hipLaunchKernelGGL((k_latent_w<true>), grid, dim3(256), xlds ? xbytes : 0, st, pb, G, sweep, xlds);
  else
    hipLaunchKernelGGL((k_latent_w<false>), grid, dim3(256), xlds ? xbytes : 0, st, pb, G, sweep, xlds);
  if (parts) *parts = pb.wm_update ? nlb : 0;
  return hipGetLastError();
}

static hipError_t launch_pairs_w(const DevProblem &pb, hipStream_t st)
{
  hipLaunchKernelGGL(k_pairs_w, dim3((unsigned)PAIR_BLOCKS, (unsigned)pb.nchains), dim3(256), 0, st, pb);
  return hipGetLastError();
}

static hipError_t launch_coef_w(const DevProblem &pb, int colour, bool with_prec, uint32_t sweep, hipStream_t st,
                                const Accum &ac, int batch)
{
  const int nc = colour ? pb.V - pb.n0 : pb.n0;
  if (nc > 0) {
    const size_t lds = (size_t)COEFW_THREADS * coefw_stride(pb.P) * sizeof(float);
    dim3 grid((unsigned)((nc + COEFW_THREADS - 1) / COEFW_THREADS), (unsigned)pb.nchains);
    hipLaunchKernelGGL(k_coef_w, grid, dim3(COEFW_THREADS), lds, st, pb, colour, sweep, ac, batch);
  }
  if (with_prec) return launch_pairs_w(pb, st);  // the pair sums after colour 1, as the colour-1 kernel forms them
  return hipGetLastError();
}

// X staging of either body (the inversion path at G = 4 pads X to 4R quads,
// R = ceil(NQ / 4)), plus the fallback areas of method 1
static size_t latent_vl_lds(const DevProblem &pb)
{
  const size_t x = (size_t)(4 * ((pb.NQ + 3) / 4)) * pb.P * sizeof(float4);
  return x + (pb.latent_method == 1 ? vl_fallback_lds(pb.P) : 0);
}

// colour set of a colour-tiled launch over [pb.vb, pb.ve) at lane split G (0: none)
static int ctile_set(const DevProblem &pb, int G)
{
  if (G != 4 || !vl_ctiles(pb)) return 0;
  if (pb.vb == 0 && pb.ve == pb.V) return 3;
  if (pb.vb == 0 && pb.ve == pb.n0) return 1;
  if (pb.vb == pb.n0 && pb.ve == pb.V) return 2;
  return 0;
}

// colours of set cs whose leftover tiles run as pieces
static int piece_set(const DevProblem &pb, int cs)
{
  return cs & ((pb.tl_k[0] > 1 ? 1 : 0) | (pb.tl_k[1] > 1 ? 2 : 0));
}

static hipError_t launch_latent_vl(const DevProblem &pb, int G, uint32_t sweep, hipStream_t st, int *parts,
                                   int wish, uint32_t wsweep, int *pieces)
{
  const size_t lds = latent_vl_lds(pb);
  const int vpb = 4 * (64 / G);  // voxels per block tile
  const long long tiles = ((long long)(pb.ve - pb.vb) + vpb - 1) / vpb;
  const long long per = (tiles + BSTAT_SLOTS - 1) / BSTAT_SLOTS;  // tiles per block, equal shares
  int nlb = (int)((tiles + per - 1) / per);
  const bool wmu = pb.wm_update != 0, bmf = pb.latent_method == 1;
  if (G != 1 && G != 2 && G != 4 && G != 8) return hipErrorInvalidValue;
  if (wish && (G != 4 || bmf)) return hipErrorInvalidValue;
  if (pb.xbx && (G != 4 || bmf || wmu)) return hipErrorInvalidValue;  // the BF16 X'Z form is the G = 4 body's
  // colour tiles (and pieces) of the inversion path at G = 4
  const int cs = ctile_set(pb, G);
  if (cs) nlb = vl_blocks(pb, cs);
  if (pieces) *pieces = piece_set(pb, cs);
  dim3 grid((unsigned)(nlb + wish), (unsigned)pb.nchains);
#define V_(PP)                                                                                             \
  do {                                                                                                     \
    if (bmf) {                                                                                             \
      if (wmu) hipLaunchKernelGGL((k_latent_vl<PP, true, true>), grid, dim3(256), lds, st, pb, G, sweep);   \
      else hipLaunchKernelGGL((k_latent_vl<PP, false, true>), grid, dim3(256), lds, st, pb, G, sweep);      \
    } else if (G == 4) {                                                                                   \
      if (wmu) hipLaunchKernelGGL((k_latent_vl_mfm<PP, true>), grid, dim3(256), lds, st, pb, sweep, wish, wsweep, cs); \
      else if (pb.xbx) hipLaunchKernelGGL((k_latent_vl_mfm<PP, false, true>), grid, dim3(256), lds, st, pb, sweep,    \
                                          wish, wsweep, cs);                                                           \
      else hipLaunchKernelGGL((k_latent_vl_mfm<PP, false>), grid, dim3(256), lds, st, pb, sweep, wish, wsweep, cs);    \
    } else {                                                                                               \
      if (wmu) hipLaunchKernelGGL((k_latent_vl<PP, true, false>), grid, dim3(256), lds, st, pb, G, sweep);  \
      else hipLaunchKernelGGL((k_latent_vl<PP, false, false>), grid, dim3(256), lds, st, pb, G, sweep);     \
    }                                                                                                      \
  } while (0)
  BSG_DISPATCH_P(pb.P, V_)
#undef V_
  if (parts) *parts = wmu ? nlb : 0;
  return hipGetLastError();
}

// The Z-free latent launch can carry the previous sweep's Wishart draw (block
// 0 of k_latent_vl_mfm): inversion draws at four lanes per voxel.
bool latent_carries_wishart(const DevProblem &pb)
{
  const int G = pb.lat_g > 0 ? pb.lat_g : latent_vl_groups(pb);
  return pb.P <= MAX_P && G == 4 && pb.latent_method == 0 && latent_vl_lds(pb) <= 65536;
}

// Voxel ranges: latent and X'Z launches cover the sorted voxels [v0, v1), so a
// sweep can draw one colour while the other colour's coefficients update.
static bool set_range(DevProblem &pb, int v0, int v1)
{
  if (v0 < 0 || v1 > pb.V || v0 > v1) return false;
  pb.vb = v0;
  pb.ve = v1;
  return true;
}

hipError_t launch_latent(const DevProblem &pb0, bool fused, int zstore, uint32_t sweep, hipStream_t st, int v0, int v1,
                         int *bstat_parts, int wish, uint32_t wsweep, int *pieces)
{
  DevProblem pb = pb0;
  if (bstat_parts) *bstat_parts = 0;
  if (pieces) *pieces = 0;
  if (!set_range(pb, v0, v1)) return hipErrorInvalidValue;
  if (pb.P > MAX_P) {  // wide problems: the Z-free fused form over every voxel
    if (!fused || zstore != 2 || wish || v0 != 0 || v1 != pb.V || pb.latent_method != 0) return hipErrorInvalidValue;
    return launch_latent_w(pb, sweep, st, bstat_parts);
  }
  if (v0 == v1 && !wish) return hipSuccess;
  if (fused && zstore == 2) {
    // the Z-free form needs X in LDS; the WM statistic needs the whole voxel range
    if (latent_vl_lds(pb) > 65536) return hipErrorInvalidValue;
    if (pb.wm_update && (v0 != 0 || v1 != pb.V)) return hipErrorInvalidValue;
    return launch_latent_vl(pb, pb.lat_g > 0 ? pb.lat_g : latent_vl_groups(pb), sweep, st, bstat_parts, wish, wsweep,
                            pieces);
  }
  if (wish) return hipErrorInvalidValue;
  int lshift, reps;
  lane_grouping(pb, lshift, reps);
  const int vpw = 64 >> lshift;
  const long long groups = ((long long)(v1 - v0) + vpw - 1) / vpw;
  const int viter = LATENT_VITER;
  const long long waves = (groups + viter - 1) / viter;
  dim3 grid((unsigned)((waves + 3) / 4), (unsigned)pb.nchains);
  const bool xreg = reps == 1;
  if (!fused && zstore == 2) return hipErrorInvalidValue;
#define LK(PP, XR, FU, ZZ) \
  hipLaunchKernelGGL((k_latent<PP, XR, FU, ZZ>), grid, dim3(256), 0, st, pb, lshift, reps, viter, sweep)
#define L_(PP)                                                          \
  do {                                                                  \
    if (fused) {                                                        \
      if (xreg) {                                                       \
        if (zstore == 0) LK(PP, true, true, 0);                         \
        else if (zstore == 1) LK(PP, true, true, 1);                    \
        else LK(PP, true, true, 2);                                     \
      } else {                                                          \
        if (zstore == 0) LK(PP, false, true, 0);                        \
        else if (zstore == 1) LK(PP, false, true, 1);                   \
        else LK(PP, false, true, 2);                                    \
      }                                                                 \
    } else {                                                            \
      if (xreg) {                                                       \
        if (zstore == 0) LK(PP, true, false, 0);                        \
        else LK(PP, true, false, 1);                                    \
      } else {                                                          \
        if (zstore == 0) LK(PP, false, false, 0);                       \
        else LK(PP, false, false, 1);                                   \
      }                                                                 \
    }                                                                   \
  } while (0)
  BSG_DISPATCH_P(pb.P, L_)
#undef L_
#undef LK
  return hipGetLastError();
}

hipError_t launch_bwm(const DevProblem &pb, int parts, uint32_t sweep, hipStream_t st)
{
  if (!pb.wm_update) return hipSuccess;
  if (parts <= 0 && pb.P > MAX_P) return hipErrorInvalidValue;  // the wide latent kernel forms the partials
  if (parts <= 0) {
    if (pb.Z == nullptr || pb.NQ <= 0) return hipErrorInvalidValue;
    parts = std::min(BSTAT_SLOTS, (pb.V + 255) / 256);
    dim3 grid((unsigned)parts, (unsigned)pb.nchains);
#define B_(PP) hipLaunchKernelGGL(k_bstat<PP>, grid, dim3(256), 0, st, pb)
    BSG_DISPATCH_P(pb.P, B_)
#undef B_
  }
  hipLaunchKernelGGL(k_bwm, dim3(1, (unsigned)pb.nchains), dim3(256), 0, st, pb, parts, sweep);
  return hipGetLastError();
}

hipError_t launch_xtz(const DevProblem &pb0, int mode, hipStream_t st, int v0, int v1)
{
  DevProblem pb = pb0;
  if (!set_range(pb, v0, v1) || pb.P > MAX_P) return hipErrorInvalidValue;
  if (v0 == v1) return hipSuccess;
  if (mode == 1) {
    const long long tiles = ((long long)(v1 - v0) + 15) / 16;
    dim3 grid((unsigned)((tiles + 3) / 4), (unsigned)pb.nchains);
#define M_(PP) hipLaunchKernelGGL(k_xtz_mfma<PP>, grid, dim3(256), 0, st, pb)
    BSG_DISPATCH_P(pb.P, M_)
#undef M_
    return hipGetLastError();
  }
  int lshift, reps;
  lane_grouping(pb, lshift, reps);
  const int vpw = 64 >> lshift;
  const long long groups = ((long long)(v1 - v0) + vpw - 1) / vpw;
  const int viter = 8;
  const long long waves = (groups + viter - 1) / viter;
  dim3 grid((unsigned)((waves + 3) / 4), (unsigned)pb.nchains);
#define L_(PP)                                                                                  \
  do {                                                                                          \
    if (reps == 1)                                                                              \
      hipLaunchKernelGGL((k_xtz_f32<PP, true>), grid, dim3(256), 0, st, pb, lshift, reps, viter);  \
    else                                                                                        \
      hipLaunchKernelGGL((k_xtz_f32<PP, false>), grid, dim3(256), 0, st, pb, lshift, reps, viter); \
  } while (0)
  BSG_DISPATCH_P(pb.P, L_)
#undef L_
  return hipGetLastError();
}

hipError_t launch_coef(const DevProblem &pb, int colour, bool with_prec, uint32_t sweep, hipStream_t st, int tail,
                       const Accum *ac, int batch)
{
  const Accum acc = ac ? *ac : Accum{};
  const int nc = colour ? pb.V - pb.n0 : pb.n0;
  if (pb.P > MAX_P) return launch_coef_w(pb, colour, with_prec, sweep, st, acc, batch);
  if (nc == 0) return with_prec ? hipErrorInvalidValue : hipSuccess;
  const int extra = (colour == 0 && (tail & 2)) ? vl_range(pb, 1).left : 0;
  dim3 grid((unsigned)((nc + 255) / 256 + extra), (unsigned)pb.nchains);
#define L_(PP)                                                                                \
  do {                                                                                        \
    if (with_prec)                                                                            \
      hipLaunchKernelGGL((k_coef<PP, true>), grid, dim3(256), 0, st, pb, colour, sweep, tail, acc, batch);  \
    else                                                                                      \
      hipLaunchKernelGGL((k_coef<PP, false>), grid, dim3(256), 0, st, pb, colour, sweep, tail, acc, batch); \
  } while (0)
  BSG_DISPATCH_P(pb.P, L_)
#undef L_
  return hipGetLastError();
}

// The overlapped sweep: Z-free inversion draws at four lanes per voxel, b
// fixed, Lambda drawn inside phase 0 (as latent_carries_wishart), both
// colours present; any number of chains (the grid's second dimension; one
// chain runs colour tiles and pieces). Phase 1 draws the Bartlett factor of
// its sweep's Lambda ahead, for the phase 0 after it.
bool overlap_supported(const DevProblem &pb)
{
  return pb.P <= MAX_P && pb.zfree && pb.latent_method == 0 && !pb.wm_update && vl_groups(pb) == 4 && pb.n0 > 0 &&
         pb.V > pb.n0 &&
         latent_carries_wishart(pb) && pb.lamtag != nullptr;
}

hipError_t launch_overlap(const DevProblem &pb, int phase, uint32_t sweep, int lat, int wish, uint32_t wsweep,
                          int tail, hipStream_t st, const Accum *ac, int batch)
{
  if (!overlap_supported(pb) || (phase != 0 && phase != 1) || (phase == 1 && wish)) return hipErrorInvalidValue;
  const Accum acc = ac ? *ac : Accum{};
  const int colour = phase ? 1 : 0;  // the updated colour; the other one is drawn
  const VlRange rd = vl_range(pb, phase ? 0 : 1);  // the drawn colour
  const int nlat = (phase == 0 || lat) ? (vl_ctiles(pb) ? vl_blocks(pb, phase ? 1 : 2) : (rd.end - rd.base + 63) / 64) : 0;
  const int ncb = ((colour ? pb.V - pb.n0 : pb.n0) + 255) / 256;
  const size_t lds = latent_vl_lds(pb);
  dim3 grid((unsigned)(wish + nlat + ncb + (phase ? 1 : 0)), (unsigned)pb.nchains);
#define L_(PP)                                                                                                     \
  do {                                                                                                             \
    if (vl_ctiles(pb) && pb.xbx)                                                                                   \
      hipLaunchKernelGGL((k_overlap<PP, false, true>), grid, dim3(256), lds, st, pb, phase, sweep, nlat, wish,     \
                         wsweep, tail, acc, batch);                                                                \
    else if (vl_ctiles(pb))                                                                                        \
      hipLaunchKernelGGL((k_overlap<PP, false>), grid, dim3(256), lds, st, pb, phase, sweep, nlat, wish, wsweep,   \
                         tail, acc, batch);                                                                        \
    else if (pb.xbx)                                                                                               \
      hipLaunchKernelGGL((k_overlap<PP, true, true>), grid, dim3(256), lds, st, pb, phase, sweep, nlat, wish,      \
                         wsweep, tail, acc, batch);                                                                \
    else                                                                                                           \
      hipLaunchKernelGGL((k_overlap<PP, true>), grid, dim3(256), lds, st, pb, phase, sweep, nlat, wish, wsweep,    \
                         tail, acc, batch);                                                                        \
  } while (0)
  BSG_DISPATCH_P(pb.P, L_)
#undef L_
  return hipGetLastError();
}

bool persist_supported(const DevProblem &pb)
{
  return overlap_supported(pb) && pb.P <= 6 && pb.nchains == 1 && vl_ctiles(pb) && !pb.xbx;
}

int persist_max_blocks(const DevProblem &pb)
{
  int per_cu = 0;
  const size_t lds = latent_vl_lds(pb);
#define L_(PP) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_persist<PP>, 256, lds)
  switch (pb.P) {
  case 1: L_(1); break;
  case 2: L_(2); break;
  case 3: L_(3); break;
  case 4: L_(4); break;
  case 5: L_(5); break;
  case 6: L_(6); break;
  default: return 0;
  }
#undef L_
  return per_cu * pb.ncu;
}

hipError_t launch_persist(const DevProblem &pb, const PTags &tg, int nblocks, hipStream_t st)
{
  if (!persist_supported(pb) || nblocks < 1 || tg.K < 1 || tg.K > PERSIST_MAX_K || tg.plen < 1 || tg.windows < tg.K ||
      tg.epoch < 1 || tg.epoch > (uint32_t)PERSIST_MAX_EPOCH)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_persist_init, dim3(1), dim3(64), 0, st, tg.head);
  const size_t lds = latent_vl_lds(pb);
  const PArgs args{pb, tg};
#define L_(PP) hipLaunchKernelGGL(k_persist<PP>, dim3((unsigned)nblocks), dim3(256), lds, st, args)
  switch (pb.P) {
  case 1: L_(1); break;
  case 2: L_(2); break;
  case 3: L_(3); break;
  case 4: L_(4); break;
  case 5: L_(5); break;
  case 6: L_(6); break;
  default: return hipErrorInvalidValue;
  }
#undef L_
  return hipGetLastError();
}

// Sampler probe: the draws of the latent kernels without their data. Each
// thread has four means spread over [-4, 2], moved by a step that changes
// every round (so that nothing of the draw is the same from round to round),
// and a response per subject (about one in eight a lesion); it draws `rounds`
// quads from its own Philox counters, summing the draws so that none is
// dropped. Its draw rate is the ceiling of the latent kernels, which add the
// means (MFMA), X'Z, the loads of b_j and y and the stores of m_j.
__global__ __launch_bounds__(256) void k_sampler_probe(float *out, int rounds)
{
  const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
  uint32_t h[4];
  philox10(gid, 0u, TAG_PROBE, 1u, 0x5A3F1C27u, 0x0BADCAFEu, h);
  float mu[4];
#pragma unroll
  for (int l = 0; l < 4; l++) mu[l] = fmaf(6.0f, u01f(h[l]), -4.0f);
  const uint32_t yq = ((h[0] & 7u) == 0u ? 1u : 0u) | ((h[1] & 7u) == 0u ? 2u : 0u) |
                      ((h[2] & 7u) == 0u ? 4u : 0u) | ((h[3] & 7u) == 0u ? 8u : 0u);
  float acc = 0.0f;
  for (int r = 0; r < rounds; r++) {
    uint32_t rr[4];
    philox10(gid, (uint32_t)r, TAG_PROBE, 0u, 0x5A3F1C27u, 0x0BADCAFEu, rr);
    const float dr = (float)(r & 31) * 0.0625f - 1.0f;
    const float m[4] = {mu[0] + dr, mu[1] + dr, mu[2] + dr, mu[3] + dr};
    float z[4];
    latent_quad(m, yq, rr, 1e-6f, z);
    acc += (z[0] + z[1]) + (z[2] + z[3]);
  }
  if (acc == 1234.5f) out[blockIdx.x] = acc;
}

hipError_t launch_sampler_probe(float *out, int blocks, int rounds, hipStream_t st)
{
  hipLaunchKernelGGL(k_sampler_probe, dim3((unsigned)blocks), dim3(256), 0, st, out, rounds);
  return hipGetLastError();
}

hipError_t launch_zero64(unsigned long long *p, size_t n, hipStream_t st)
{
  if (n == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_zero64, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_pairs(const DevProblem &pb, hipStream_t st)
{
  if (pb.P > MAX_P) return launch_pairs_w(pb, st);
  dim3 grid((unsigned)PAIR_BLOCKS, (unsigned)pb.nchains);
#define L_(PP) hipLaunchKernelGGL(k_pairs<PP>, grid, dim3(256), 0, st, pb)
  BSG_DISPATCH_P(pb.P, L_)
#undef L_
  return hipGetLastError();
}

hipError_t launch_wishart(const DevProblem &pb, bool side, uint32_t sweep, hipStream_t st)
{
  dim3 grid((unsigned)pb.nchains);
  if (pb.P > MAX_P) {  // always from the PAIR_BLOCKS partials of k_pairs_w
    hipLaunchKernelGGL(k_wishart_w, grid, dim3(256), 0, st, pb, sweep);
    return hipGetLastError();
  }
#define L_(PP)                                                                   \
  do {                                                                           \
    if (side)                                                                    \
      hipLaunchKernelGGL((k_wishart<PP, true>), grid, dim3(256), 0, st, pb, sweep);     \
    else                                                                         \
      hipLaunchKernelGGL((k_wishart<PP, false>), grid, dim3(256), 0, st, pb, sweep);    \
  } while (0)
  BSG_DISPATCH_P(pb.P, L_)
#undef L_
  return hipGetLastError();
}

hipError_t launch_cmeans(const DevProblem &pb, double *part, int blocks, hipStream_t st)
{
  dim3 grid((unsigned)blocks, (unsigned)pb.nchains);
  if (pb.P > MAX_P) {
    hipLaunchKernelGGL(k_cmeans_w, grid, dim3(256), 0, st, pb, part);
    return hipGetLastError();
  }
#define L_(PP) hipLaunchKernelGGL(k_cmeans<PP>, grid, dim3(256), 0, st, pb, part)
  BSG_DISPATCH_P(pb.P, L_)
#undef L_
  return hipGetLastError();
}

int loglik_tiles(const DevProblem &pb) { return (pb.V + 255) / 256; }

hipError_t launch_loglik(const DevProblem &pb, int T, double *part, double *out, hipStream_t st)
{
  if (T < 1 || T > pb.P) return hipErrorInvalidValue;
  const int tiles = loglik_tiles(pb);
  dim3 grid((unsigned)tiles, (unsigned)((pb.N + 255) / 256), (unsigned)pb.nchains);
  if (pb.P > MAX_P) {
    hipLaunchKernelGGL(k_loglik_w, grid, dim3(256), 0, st, pb, T, part);
  } else {
#define L_(PP) hipLaunchKernelGGL(k_loglik<PP>, grid, dim3(256), 0, st, pb, T, part)
    BSG_DISPATCH_P(pb.P, L_)
#undef L_
  }
  const int NT = pb.N * T;
  hipLaunchKernelGGL(k_llreduce, dim3((unsigned)((NT + 255) / 256), (unsigned)pb.nchains), dim3(256), 0, st, tiles,
                     NT, (const double *)part, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Allocation-free scratch (bsg_coef_means_noalloc): a fixed device array for
// callers that cannot allocate. Each device holds its own copy until the
// process ends. Host code never passes the array's name to a kernel (on the
// host it is only a stand-in): scratch_init takes the array's address on the
// current device once through hipGetSymbolAddress, and kernels receive that
// address. Every call on a device shares the array, so calls that overlap on
// one device race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[MAX_DEVICES] = {};

hipError_t scratch_init()
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return hipSuccess;
  void *p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[dev] = static_cast<double *>(p);
  return hipSuccess;
}

double *scratch_ptr()
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[dev];
}

// out[c] = sum over b < blocks of part[b P + c], in block order (the order the
// host adds k_cmeans's partials in bsg_coef_means)
__global__ __launch_bounds__(64) void k_cmeans_fold(const double *part, int blocks, int P, double *out)
{
  const int c = threadIdx.x;
  if (c >= P) return;
  double acc = 0.0;
  for (int b = 0; b < blocks; b++) acc += part[(size_t)b * P + c];
  out[c] = acc;
}

hipError_t launch_cmeans_fold(const double *part, int blocks, int P, double *out, hipStream_t st)
{
  if (P < 1 || P > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_cmeans_fold, dim3(1), dim3(64), 0, st, part, blocks, P, out);
  return hipGetLastError();
}

// Advances the device sweep base at the end of a captured group of sweeps.
__global__ void k_tick(uint32_t *base, uint32_t k)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *base += k;
}

hipError_t launch_tick(uint32_t *base, uint32_t k, hipStream_t st)
{
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, st, base, k);
  return hipGetLastError();
}

hipError_t launch_latent_multi(int P, const DevProblem *pbs, const WorkItem *items, int n, size_t lds, bool bmf,
                               bool g4, uint32_t sweep, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  if (lds > 65536 || (g4 && bmf)) return hipErrorInvalidValue;
  const dim3 grid((unsigned)n);
#define L_(PP)                                                                                             \
  do {                                                                                                     \
    if (g4) hipLaunchKernelGGL((k_latent_multi_mfm<PP>), grid, dim3(256), lds, st, pbs, items, sweep);       \
    else if (bmf) hipLaunchKernelGGL((k_latent_multi<PP, true>), grid, dim3(256), lds, st, pbs, items, sweep); \
    else hipLaunchKernelGGL((k_latent_multi<PP, false>), grid, dim3(256), lds, st, pbs, items, sweep);       \
  } while (0)
  BSG_DISPATCH_P(P, L_)
#undef L_
  return hipGetLastError();
}

hipError_t launch_latent_multi_z(int P, const DevProblem *pbs, const WorkItem *items, int n, uint32_t sweep,
                                 hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  if (P < 1 || P > MAX_P) return hipErrorInvalidValue;
#define L_(PP) hipLaunchKernelGGL(k_latent_multi_z<PP>, dim3((unsigned)n), dim3(256), 0, st, pbs, items, sweep)
  BSG_DISPATCH_P(P, L_)
#undef L_
  return hipGetLastError();
}

hipError_t launch_xtz_multi(int P, const DevProblem *pbs, const WorkItem *items, int n, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  if (P < 1 || P > MAX_P) return hipErrorInvalidValue;
#define L_(PP) hipLaunchKernelGGL(k_xtz_mfma_multi<PP>, dim3((unsigned)n), dim3(256), 0, st, pbs, items)
  BSG_DISPATCH_P(P, L_)
#undef L_
  return hipGetLastError();
}

hipError_t launch_coef_multi(int P, const DevProblem *pbs, const WorkItem *items, int n, int colour, bool prec,
                             uint32_t sweep, hipStream_t st, const Accum *accs, const BatchIndex &bidx)
{
  if (n <= 0) return hipSuccess;
#define L_(PP)                                                                                            \
  do {                                                                                                    \
    if (prec)                                                                                             \
      hipLaunchKernelGGL((k_coef_multi<PP, true>), dim3((unsigned)n), dim3(256), 0, st, pbs, items, colour, sweep, \
                         accs, bidx);                                                                     \
    else                                                                                                  \
      hipLaunchKernelGGL((k_coef_multi<PP, false>), dim3((unsigned)n), dim3(256), 0, st, pbs, items, colour, sweep, \
                         accs, bidx);                                                                     \
  } while (0)
  BSG_DISPATCH_P(P, L_)
#undef L_
  return hipGetLastError();
}

hipError_t launch_wishart_multi(int P, const DevProblem *pbs, const WorkItem *items, int n, uint32_t sweep,
                                hipStream_t st)
{
  if (n <= 0) return hipSuccess;
#define L_(PP) hipLaunchKernelGGL(k_wishart_multi<PP>, dim3((unsigned)n), dim3(256), 0, st, pbs, items, sweep)
  BSG_DISPATCH_P(P, L_)
#undef L_
  return hipGetLastError();
}


int latent_vl_tiles(const DevProblem &pb, int G) { return (pb.V + 4 * (64 / G) - 1) / (4 * (64 / G)); }

int latent_vl_blocks_g4(const DevProblem &pb)
{
  return vl_ctiles(pb) ? vl_blocks(pb, 3) : (pb.V + 63) / 64;
}

size_t latent_vl_lds_bytes(const DevProblem &pb) { return latent_vl_lds(pb); }

hipError_t launch_accum(const DevProblem &pb, const Accum &ac, int batch, hipStream_t st)
{
  dim3 grid((unsigned)((pb.V + 255) / 256), (unsigned)pb.nchains);
  if (pb.P > MAX_P) {
    hipLaunchKernelGGL(k_accum_w, grid, dim3(256), 0, st, pb, ac, batch);
    return hipGetLastError();
  }
#define L_(PP) hipLaunchKernelGGL(k_accum<PP>, grid, dim3(256), 0, st, pb, ac, batch)
  BSG_DISPATCH_P(pb.P, L_)
#undef L_
  return hipGetLastError();
}

}  // namespace bsg
