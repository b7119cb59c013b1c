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

namespace bsg {

// ---------------------------------------------------------------------------
// Philox4x32-10 and uniforms
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                         uint32_t k0, uint32_t k1, uint32_t r[4])
{
#pragma unroll
  for (int i = 0; i < 10; i++) {
    uint32_t hi0 = __umulhi(0xD2511F53u, c0);
    uint32_t lo0 = 0xD2511F53u * c0;
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
    uint32_t lo1 = 0xCD9E8D57u * c2;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  r[0] = c0;
  r[1] = c1;
  r[2] = c2;
  r[3] = c3;
}

// odd multiples of 2^-24: exact in float and double, never 0 or 1
__device__ __forceinline__ float u01f(uint32_t x)
{
  return (float)(((x >> 9) << 1) | 1u) * 5.9604644775390625e-8f;
}
__device__ __forceinline__ double u01d(uint32_t x)
{
  return (double)(((x >> 9) << 1) | 1u) * 5.9604644775390625e-8;
}

// ---------------------------------------------------------------------------
// Truncated-normal latent draw by inversion (one uniform, no rejection loop).
// W ~ N(0,1) truncated below at a; Z = mu + s W with s = +1 (y=1), -1 (y=0)
// and a = eps - s mu. The upper limit 500 (covarGPU.cu:1415,1418) has
// Q(500 - s mu) = 0 in FP32 for |mu| < 480 and is not evaluated.
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

__device__ __forceinline__ float latent_draw(float mu, uint32_t y, float u, float eps)
{
  float s = y ? 1.0f : -1.0f;
  float a = eps - s * mu;
  float w;
  if (a >= 8.0f) {
    w = tail_w(a, u);
  } else {
    float qt = 0.5f * erfcf(fabsf(a) * 0.70710678118654752f);  // Q(|a|) <= 1/2
    float arg, sg = 1.0f;
    if (a >= 0.0f) {
      arg = u * qt;
    } else {
      float p = u * (1.0f - qt);
      if (p <= 0.5f) {
        arg = p;
      } else {
        arg = (1.0f - u) + u * qt;  // 1 - p without cancellation
        sg = -1.0f;
      }
    }
    w = sg * 1.41421356237309505f * erfcinvf(2.0f * arg);
  }
  w = fmaxf(w, a);
  return mu + s * w;
}

// ---------------------------------------------------------------------------
// k_latent. Lanes of a wave are split into groups of L = 2^lshift lanes; a
// group owns one voxel and lane ql of the group owns the subject quads
// q = ql, ql + L, ... (4 subjects per Philox block, one float4 store).
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(256) void k_latent(DevProblem pb, int lshift, int reps)
{
  const int chain = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int s = (wave << (6 - lshift)) + (lane >> lshift);
  if (s >= pb.V) return;
  const int ql = lane & ((1 << lshift) - 1);
  const uint32_t sweep = pb.sweep[chain];
  const float b = pb.bwm[chain];
  const float *bet = pb.beta + ((size_t)chain * pb.V + s) * P;
  float bs[P];
#pragma unroll
  for (int c = 0; c < P; c++) bs[c] = bet[c];
  const float base = b * pb.wm[s];
  float *zrow = pb.Z + ((size_t)chain * pb.V + s) * pb.NS;
  const uint32_t *yrow = pb.ybits + (size_t)s * pb.NW;
  for (int r = 0; r < reps; r++) {
    const int q = ql + (r << lshift);
    if (q >= pb.NQ) break;
    const int i0 = 4 * q;
    float mu[4] = {base, base, base, base};
#pragma unroll
    for (int c = 0; c < P; c++) {
      float4 x = *reinterpret_cast<const float4 *>(pb.X + (size_t)c * pb.NS + i0);
      mu[0] = fmaf(bs[c], x.x, mu[0]);
      mu[1] = fmaf(bs[c], x.y, mu[1]);
      mu[2] = fmaf(bs[c], x.z, mu[2]);
      mu[3] = fmaf(bs[c], x.w, mu[3]);
    }
    const uint32_t yw = yrow[q >> 3] >> ((q & 7) * 4);
    uint32_t rr[4];
    philox10((uint32_t)s * (uint32_t)pb.NQ + (uint32_t)q, sweep, TAG_LATENT, (uint32_t)chain, pb.key0,
             pb.key1, rr);
    float z[4];
#pragma unroll
    for (int l = 0; l < 4; l++)
      z[l] = (i0 + l < pb.N) ? latent_draw(mu[l], (yw >> l) & 1u, u01f(rr[l]), pb.trunc_eps) : 0.0f;
    *reinterpret_cast<float4 *>(zrow + i0) = make_float4(z[0], z[1], z[2], z[3]);
  }
}

// ---------------------------------------------------------------------------
// k_xtz: m_j = X'Z_j - b WM_j X'1, FP32 FMA, same lane grouping as k_latent.
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(256) void k_xtz_f32(DevProblem pb, int lshift, int reps)
{
  const int chain = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int s = (wave << (6 - lshift)) + (lane >> lshift);
  const bool active = s < pb.V;
  const int ql = lane & ((1 << lshift) - 1);
  float acc[P];
#pragma unroll
  for (int c = 0; c < P; c++) acc[c] = 0.0f;
  if (active) {
    const float *zrow = pb.Z + ((size_t)chain * pb.V + s) * pb.NS;
    for (int r = 0; r < reps; r++) {
      const int q = ql + (r << lshift);
      if (q >= pb.NQ) break;
      const int i0 = 4 * q;
      float4 z = *reinterpret_cast<const float4 *>(zrow + i0);
#pragma unroll
      for (int c = 0; c < P; c++) {
        float4 x = *reinterpret_cast<const float4 *>(pb.X + (size_t)c * pb.NS + i0);
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
  if (active && ql == 0) {
    const float bw = pb.bwm[chain] * pb.wm[s];
    float *m = pb.M + ((size_t)chain * pb.V + s) * P;
#pragma unroll
    for (int c = 0; c < P; c++) m[c] = acc[c] - bw * pb.xsum[c];
  }
}

// ---------------------------------------------------------------------------
// Coefficient update. With Q = X'X + n_j Lambda and r = m + Lambda sumV, the
// reference draws b = Q^-1 r + chol_lower(Q^-1) z (covarGPU.cu:374-386).
// Factor Q = U U' with U upper (Cholesky of the index-reversed Q); then
// chol_lower(Q^-1) = U^-T and b = U^-T (U^-1 r + z): one factorisation and two
// triangular solves, no explicit inverse.
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
    T ukk = sqrt(d);
    U[k][k] = ukk;
    dinv[k] = T(1) / ukk;
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

template <int P>
__global__ __launch_bounds__(256) void k_coef(DevProblem pb, int colour)
{
  const int chain = blockIdx.y;
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int nc = colour ? pb.V - pb.n0 : pb.n0;
  if (q >= nc) return;
  const int s = q + (colour ? pb.n0 : 0);
  const uint32_t sweep = pb.sweep[chain];
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
#pragma unroll
  for (int d = 0; d < 6; d++) {
    if (nb[d] >= 0) {
      nn++;
      const float *bk = betc + (size_t)nb[d] * P;
#pragma unroll
      for (int c = 0; c < P; c++) sv[c] += bk[c];
    }
  }
  const float *m = pb.M + ((size_t)chain * pb.V + s) * P;
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
      float rad = sqrtf(-2.0f * logf(u01f(rr[2 * h])));
      float sn, cs;
      sincospif(2.0f * u01f(rr[2 * h + 1]), &sn, &cs);
      if (4 * g + 2 * h < P) z[4 * g + 2 * h] = rad * cs;
      if (4 * g + 2 * h + 1 < P) z[4 * g + 2 * h + 1] = rad * sn;
    }
  }
  float out[P];
  float *dst = pb.beta + ((size_t)chain * pb.V + s) * P;
  if (!ul_draw<P, float>(Q, r, z, out, pb.pivot_min)) {
    // FP64 refactorisation of this voxel (FP32 pivot below pivot_min)
    double Qd[P][P], rd[P], zd[P], od[P];
#pragma unroll
    for (int i = 0; i < P; i++) {
      double acc = (double)m[i];
#pragma unroll
      for (int j = 0; j < P; j++) {
        double lij = (double)lam[i * P + j];
        acc += lij * (double)sv[j];
        Qd[i][j] = (double)pb.xtx[i * P + j] + (double)nn * lij;
      }
      rd[i] = acc;
      zd[i] = (double)z[i];
    }
    atomicAdd(&pb.flags[0], 1u);
    if (!ul_draw<P, double>(Qd, rd, zd, od, 0.0)) {
      atomicAdd(&pb.flags[1], 1u);
      return;  // not positive definite: keep the previous coefficients
    }
#pragma unroll
    for (int c = 0; c < P; c++) out[c] = (float)od[c];
  }
#pragma unroll
  for (int c = 0; c < P; c++) dst[c] = out[c];
}

// ---------------------------------------------------------------------------
// k_pairs: T_ab = sum over mask voxels j and their +x,+y,+z neighbours k with
// b^a_k != 0 (covarGPU.cu:975) of (b^a_j - b^a_k)(b^b_j - b^b_k). Float
// differences as the reference, exact FP64 products, FP64 sums; one partial
// per block, fixed block count, fixed reduction order (bitwise reproducible).
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
      float bk[P], df[P];
#pragma unroll
      for (int c = 0; c < P; c++) {
        bk[c] = betc[(size_t)k * P + c];
        df[c] = bj[c] - bk[c];
      }
      int e = 0;
#pragma unroll
      for (int a = 0; a < P; a++) {
        const bool ok = bk[a] != 0.0f;
        const double da = ok ? (double)df[a] : 0.0;
#pragma unroll
        for (int c = 0; c <= a; c++, e++) acc[e] = fma(da, (double)df[c], acc[e]);
      }
    }
  }
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
    double v = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
    pb.partials[((size_t)chain * PAIR_BLOCKS + blockIdx.x) * NE + e] = v;
  }
}

// ---------------------------------------------------------------------------
// k_wishart: one wave per chain. Sums the partials, forms S = (T+I)^-1 and
// L = chol(S) in FP64 (covarGPU.cu:1086-1101), draws the Bartlett factor A
// (A_ii^2 ~ chi2(df-i) by Marsaglia-Tsang gamma, A_ij ~ N(0,1)) and writes
// Lambda = (L A)(L A)' (cholesky.cpp:259-281). Advances the chain's sweep.
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

template <int P>
__global__ __launch_bounds__(64) void k_wishart(DevProblem pb)
{
  constexpr int NE = P * (P + 1) / 2;
  const int chain = blockIdx.x;
  const int lane = threadIdx.x;
  __shared__ double T[NE];
  __shared__ double A[P][P];
  const uint32_t sweep = pb.sweep[chain];
  const double *part = pb.partials + (size_t)chain * PAIR_BLOCKS * NE;
  for (int e = 0; e < NE; e++) {
    double v = 0.0;
    for (int b = lane; b < PAIR_BLOCKS; b += 64) v += part[(size_t)b * NE + e];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) T[e] = v;
  }
  for (int idx = lane; idx < P * P; idx += 64) {
    const int i = idx / P, j = idx % P;
    double val = 0.0;
    if (i == j) {
      val = sqrt(2.0 * gamma_mt(0.5 * (double)(pb.df - i), (uint32_t)i, sweep, (uint32_t)chain,
                                pb.key0, pb.key1));
    } else if (j < i) {
      uint32_t rr[4];
      philox10((uint32_t)idx, sweep, TAG_WNORM, (uint32_t)chain, pb.key0, pb.key1, rr);
      val = sqrt(-2.0 * log(u01d(rr[0]))) * cospi(2.0 * u01d(rr[1]));
    }
    A[i][j] = val;
  }
  __syncthreads();
  if (lane != 0) return;
  double S[P][P];
  for (int a = 0, e = 0; a < P; a++)
    for (int c = 0; c <= a; c++, e++) {
      S[a][c] = T[e];
      S[c][a] = T[e];
      pb.tsum[(size_t)chain * P * P + a * P + c] = T[e];
      pb.tsum[(size_t)chain * P * P + c * P + a] = T[e];
    }
  for (int a = 0; a < P; a++) S[a][a] += 1.0;
  chol_lower<P>(S);
  // S <- (L L')^-1, column by column
  double inv[P][P];
  for (int col = 0; col < P; col++) {
    double y[P];
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
  double B[P][P];
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
      lam64[i * P + j] = v;
      lam[i * P + j] = (float)v;
    }
  pb.sweep[chain] = sweep + 1;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static void lane_grouping(const DevProblem &pb, int &lshift, int &reps)
{
  int L = 1;
  lshift = 0;
  while (L < pb.NQ && L < 64) {
    L <<= 1;
    lshift++;
  }
  reps = (pb.NQ + L - 1) / L;
}


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

hipError_t launch_latent(const DevProblem &pb, hipStream_t st)
{
  int lshift, reps;
  lane_grouping(pb, lshift, reps);
  const long long waves = ((long long)pb.V + (64 >> lshift) - 1) / (64 >> lshift);
  dim3 grid((unsigned)((waves + 3) / 4), (unsigned)pb.nchains);
#define L_(PP) hipLaunchKernelGGL(k_latent<PP>, grid, dim3(256), 0, st, pb, lshift, reps)
  BSG_DISPATCH_P(pb.P, L_)
#undef L_
  return hipGetLastError();
}

hipError_t launch_xtz(const DevProblem &pb, int mode, hipStream_t st)
{
  (void)mode;
  int lshift, reps;
  lane_grouping(pb, lshift, reps);
  const long long waves = ((long long)pb.V + (64 >> lshift) - 1) / (64 >> lshift);
  dim3 grid((unsigned)((waves + 3) / 4), (unsigned)pb.nchains);
#define L_(PP) hipLaunchKernelGGL(k_xtz_f32<PP>, grid, dim3(256), 0, st, pb, lshift, reps)
  BSG_DISPATCH_P(pb.P, L_)
#undef L_
  return hipGetLastError();
}

hipError_t launch_coef(const DevProblem &pb, int colour, hipStream_t st)
{
  const int nc = colour ? pb.V - pb.n0 : pb.n0;
  if (nc == 0) return hipSuccess;
  dim3 grid((unsigned)((nc + 255) / 256), (unsigned)pb.nchains);
#define L_(PP) hipLaunchKernelGGL(k_coef<PP>, grid, dim3(256), 0, st, pb, colour)
  BSG_DISPATCH_P(pb.P, L_)
#undef L_
  return hipGetLastError();
}

hipError_t launch_pairs(const DevProblem &pb, hipStream_t st)
{
  dim3 grid((unsigned)PAIR_BLOCKS, (unsigned)pb.nchains);
#define L_(PP) hipLaunchKernelGGL(k_pairs<PP>, grid, dim3(256), 0, st, pb)
  BSG_DISPATCH_P(pb.P, L_)
#undef L_
  return hipGetLastError();
}

hipError_t launch_wishart(const DevProblem &pb, hipStream_t st)
{
  dim3 grid((unsigned)pb.nchains);
#define L_(PP) hipLaunchKernelGGL(k_wishart<PP>, grid, dim3(64), 0, st, pb)
  BSG_DISPATCH_P(pb.P, L_)
#undef L_
  return hipGetLastError();
}

}  // namespace bsg
