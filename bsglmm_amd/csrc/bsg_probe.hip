// bsg_probe.hip -- stream kernels that calibrate bandwidth and counter figures
// (a copy, a write-only and a read-only stream over one buffer), and VALU
// kernels that measure the chip's wave64 instruction issue rate.
#include "bsg_internal.h"

namespace bsg {

__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n)
{
  const size_t stride = (size_t)gridDim.x * 256;
  size_t i = blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const float4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
    dst[i] = a;
    dst[i + stride] = b;
    dst[i + 2 * stride] = c;
    dst[i + 3 * stride] = d;
  }
  for (; i < n; i += stride) dst[i] = src[i];
}

// write-only stream: consecutive lanes store consecutive float4s, as k_latent does
__global__ __launch_bounds__(256) void k_fill(float4 *__restrict__ dst, size_t n)
{
  const size_t stride = (size_t)gridDim.x * 256;
  const float4 v = make_float4(1.f, 2.f, 3.f, (float)blockIdx.x);
  for (size_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = v;
}

// read-only stream; a data-dependent store keeps the loads alive
__global__ __launch_bounds__(256) void k_sum(const float4 *__restrict__ src, size_t n, float *out)
{
  const size_t stride = (size_t)gridDim.x * 256;
  float acc = 0.f;
  size_t i = blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const float4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
    acc += (a.x + b.y) + (c.z + d.w);
  }
  for (; i < n; i += stride) acc += src[i].x;
  if (acc == 12345.f) out[blockIdx.x] = acc;
}

hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st)
{
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, st, static_cast<const float4 *>(src),
                     static_cast<float4 *>(dst), bytes / 16);
  return hipGetLastError();
}

// dst must hold at least 8192 floats for kind 2
hipError_t launch_probe(int kind, const void *src, void *dst, size_t bytes, hipStream_t st)
{
  const size_t n = bytes / 16;
  if (kind == 0) return launch_copy(src, dst, bytes, st);
  if (kind == 1)
    hipLaunchKernelGGL(k_fill, dim3(8192), dim3(256), 0, st, static_cast<float4 *>(dst), n);
  else
    hipLaunchKernelGGL(k_sum, dim3(8192), dim3(256), 0, st, static_cast<const float4 *>(src), n,
                       static_cast<float *>(dst));
  return hipGetLastError();
}

// VALU issue probe: every thread runs `iters` rounds of VALU_CHAINS independent
// chains of one instruction kind, 0: v_fma_f32, 1: v_exp_f32 (a
// transcendental). The loop counter lives in scalar registers, so each round
// issues exactly VALU_CHAINS vector instructions per wave; a data-dependent
// store keeps the chains alive. Eight waves per SIMD hide each chain's latency.
constexpr int VALU_CHAINS = VALU_CHAINS_PER_ROUND;

template <int KIND>
__global__ __launch_bounds__(256) void k_valu(float *out, int iters, float s)
{
  float a[VALU_CHAINS];
#pragma unroll
  for (int c = 0; c < VALU_CHAINS; c++) a[c] = 1.0f + 1e-3f * (float)threadIdx.x + (float)c;
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int c = 0; c < VALU_CHAINS; c++) a[c] = KIND == 0 ? fmaf(a[c], s, 0.5f) : __builtin_amdgcn_exp2f(a[c]);
  }
  float r = 0.0f;
#pragma unroll
  for (int c = 0; c < VALU_CHAINS; c++) r += a[c];
  if (r == 1234.5f) out[blockIdx.x] = r;
}

// blocks of 256 threads; out holds at least `blocks` floats. Wave64 VALU
// instructions in the timed loop: blocks * 4 * iters * VALU_CHAINS.
hipError_t launch_valu(int kind, float *out, int blocks, int iters, hipStream_t st)
{
  if (kind == 0)
    hipLaunchKernelGGL(k_valu<0>, dim3((unsigned)blocks), dim3(256), 0, st, out, iters, 0.99999f);
  else
    hipLaunchKernelGGL(k_valu<1>, dim3((unsigned)blocks), dim3(256), 0, st, out, iters, 0.5f);
  return hipGetLastError();
}

}  // namespace bsg
