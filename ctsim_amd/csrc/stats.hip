// Herman's comparative statistics d, r, e of two images on the device
// (ImageFile::comparativeStatistics, libctsim/imagefile.cpp:185-240).
//
// Both images are float32 column-major [nx][ny] (pixel (ix, iy) at ix*ny+iy),
// the first being the reference.  Sums run in FP64 as in the reference, in a
// fixed tree order (per-lane strided partials, wave shuffles, one partial
// per block, then one block over the partials), so repeated calls give the
// same bits; against the reference's sequential sums d and r agree to about
// 1e-15 relative.  e (the largest 2x2-block-mean difference) forms each block
// mean with the reference's float expression and order (and the pixel
// difference in float, as the reference does), so e is bit-identical.
#include <hip/hip_runtime.h>

#include "ctsim_internal.h"

namespace ctsim {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Reduces K per-thread values over the block; thread 0 writes them to out.
template <int K, bool MAX>
__device__ inline void block_reduce(double (&v)[K], double* out) {
  __shared__ double part[kThreads / 64][K];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = MAX ? wave_max(v[k]) : wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; k++) part[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; k++) {
      double s = part[0][k];
      for (int w = 1; w < kThreads / 64; w++) s = MAX ? fmax(s, part[w][k]) : s + part[w][k];
      out[k] = s;
    }
}

// Pass 1: per-block sums of the reference image.
__global__ void __launch_bounds__(kThreads) sum_kernel(const float* __restrict__ a, size_t n,
                                                       double* __restrict__ partial) {
  double v[1] = {0.0};
  for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads)
    v[0] += a[i];
  block_reduce<1, false>(v, partial + blockIdx.x);
}

// Pass 2: per-block sums of diff^2, |diff|, (a-mean)^2, |a|, reading the
// mean from the finished pass-1 reduction.
__global__ void __launch_bounds__(kThreads) moments_kernel(const float* __restrict__ a,
                                                           const float* __restrict__ b, size_t n,
                                                           const double* __restrict__ total,
                                                           double* __restrict__ partial) {
  const double mean = total[0] / (double)n;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
    const double x = a[i];
    const double diff = a[i] - b[i];  // float difference, as the reference forms it
    v[0] += diff * diff;
    v[1] += fabs(diff);
    const double dm = x - mean;
    v[2] += dm * dm;
    v[3] += fabs(x);
  }
  block_reduce<4, false>(v, partial + 4 * blockIdx.x);
}

// Pass 3: per-block maximum of |mean2x2(a) - mean2x2(b)| over the hx*hy blocks.
__global__ void __launch_bounds__(kThreads) worst_kernel(const float* __restrict__ a,
                                                         const float* __restrict__ b, int ny,
                                                         int hx, int hy, double* __restrict__ partial) {
  double v[1] = {-1.0};
  const size_t nb = (size_t)hx * hy;
  for (size_t k = blockIdx.x * (size_t)kThreads + threadIdx.x; k < nb; k += (size_t)gridDim.x * kThreads) {
    const size_t ix = k / hy, iy = k % hy;
    const size_t p00 = 2 * ix * ny + 2 * iy, p10 = p00 + ny;
    // float sums in the reference's order, then scaled in double
    const double ma = 0.25 * (double)(((a[p00] + a[p10]) + a[p00 + 1]) + a[p10 + 1]);
    const double mb = 0.25 * (double)(((b[p00] + b[p10]) + b[p00 + 1]) + b[p10 + 1]);
    v[0] = fmax(v[0], fabs(ma - mb));
  }
  block_reduce<1, true>(v, partial + blockIdx.x);
}

// Final pass: one block over K-wide partials of nPart blocks.
template <int K, bool MAX>
__global__ void __launch_bounds__(kThreads) final_kernel(const double* __restrict__ partial,
                                                         int nPart, double* __restrict__ out) {
  double v[K];
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = MAX ? -1.0 : 0.0;
  for (int i = threadIdx.x; i < nPart; i += kThreads)
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = MAX ? fmax(v[k], partial[K * i + k]) : v[k] + partial[K * i + k];
  block_reduce<K, MAX>(v, out);
}

int blocks_for(size_t n) {
  size_t b = (n + kThreads - 1) / kThreads;
  if (b < 1) b = 1;
  return (int)(b < (size_t)kMaxBlocks ? b : (size_t)kMaxBlocks);
}

}  // namespace
}  // namespace ctsim

using namespace ctsim;

extern "C" int ctsim_image_stats(ctsim_ctx* ctx, const float* d_ref, const float* d_cmp, int nx,
                                 int ny, double* h_dre) {
  if (!ctx || !d_ref || !d_cmp || !h_dre)
    return fail(CTSIM_ERR_INVALID, "ctsim_image_stats: NULL argument");
  if (nx < 1 || ny < 1) return fail(CTSIM_ERR_INVALID, "ctsim_image_stats: empty image");
  CTSIM_CHECK(ctx->activate());
  const size_t n = (size_t)nx * ny;
  const int hx = nx / 2, hy = ny / 2;
  const int b1 = blocks_for(n), b3 = blocks_for((size_t)hx * hy);
  // scratch: [b1 sums][4*b1 moments][b3 maxima][8 results]
  const size_t off2 = b1, off3 = off2 + 4 * (size_t)b1, offR = off3 + b3;
  CTSIM_CHECK(ctx->scratch.reserve((offR + 8) * sizeof(double)));
  double* s = ctx->scratch.as<double>();
  hipStream_t st = ctx->stream;
  sum_kernel<<<b1, kThreads, 0, st>>>(d_ref, n, s);
  final_kernel<1, false><<<1, kThreads, 0, st>>>(s, b1, s + offR);
  moments_kernel<<<b1, kThreads, 0, st>>>(d_ref, d_cmp, n, s + offR, s + off2);
  final_kernel<4, false><<<1, kThreads, 0, st>>>(s + off2, b1, s + offR + 1);
  if (hx > 0 && hy > 0) {
    worst_kernel<<<b3, kThreads, 0, st>>>(d_ref, d_cmp, ny, hx, hy, s + off3);
    final_kernel<1, true><<<1, kThreads, 0, st>>>(s + off3, b3, s + offR + 5);
  }
  CTSIM_HIP(hipGetLastError());
  double r[6] = {0, 0, 0, 0, 0, -1.0};
  CTSIM_HIP(hipMemcpyAsync(r, s + offR, (hx > 0 && hy > 0 ? 6 : 5) * sizeof(double),
                           hipMemcpyDeviceToHost, st));
  CTSIM_HIP(hipStreamSynchronize(st));
  h_dre[0] = sqrt(r[1] / r[3]);
  h_dre[1] = r[2] / r[4];
  h_dre[2] = r[5];
  return CTSIM_OK;
}
