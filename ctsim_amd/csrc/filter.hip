// Projection filtering kernels for gfx950 (ProcessSignal::filterSignal,
// libctsim/procsignal.cpp:544-673).
//
// Sequential-sum kernels (FP64 convolution and the exact DFT of the
// "fourier" / "fouier-table" methods): lanes run over views, so every lane of
// a wave needs the same filter tap or twiddle at the same step (an LDS
// broadcast or a scalar load) while its own sum runs in the reference's order.
//
// FFT kernels (the "fft" method, and every Fourier method in mixed
// precision): one workgroup filters two views with one complex radix-2
// Stockham transform in LDS: z = x_a + i x_b, Z = FFT(z), split Z into the two
// real spectra by Hermitian symmetry, apply the method's spectral multipliers,
// recombine, inverse FFT; the real part is view a, the imaginary part view b.
#include <hip/hip_runtime.h>

#include "filter_plan.h"
#include "fft_device.h"
#include "mixed_device.h"

#pragma clang fp contract(off)

namespace ctsim {
namespace {

constexpr int kLanes = 64;        // views per wave
constexpr int kWaves = 4;         // waves per block
constexpr int kOutPerWave = 8;    // outputs per wave (independent sums)
constexpr int kJChunk = 64;       // inputs staged per step
constexpr int kConvOut = 16;      // exact convolution: outputs per lane (independent sums)
constexpr int kKChunk = 32;       // spectrum entries staged per step (inverse DFT)

// ---------------------------------------------------- exact convolution
// out[v][i] = dx * sum_{j=0}^{n-1} x[v][j] * h[i - j + n - 1]  (procsignal.cpp:659-673)
template <class TO>
__global__ __launch_bounds__(kLanes * kWaves) void conv_seq_kernel(
    const float* __restrict__ in, const double* __restrict__ w, const double* __restrict__ h,
    TO* __restrict__ out, int nView, int n, double dx, int outStride) {
  __shared__ double s_x[kJChunk][kLanes];
  extern __shared__ double s_h[];  // the block's taps: h[ib + k], k < n + kWaves * kConvOut - 1
  const int lane = threadIdx.x & (kLanes - 1), wave = threadIdx.x / kLanes;
  const int v = blockIdx.y * kLanes + lane;
  const int ib = blockIdx.x * kWaves * kConvOut, i0 = ib + wave * kConvOut;
  // output i of the block at input j takes tap i - j + n - 1, in [ib, ib + n - 1 + kWaves kConvOut)
  const int nTap = n + kWaves * kConvOut - 1;
  for (int k = threadIdx.x; k < nTap; k += blockDim.x) s_h[k] = h[min(ib + k, 2 * n - 2)];
  double acc[kConvOut];
#pragma unroll
  for (int o = 0; o < kConvOut; o++) acc[o] = 0.0;
  for (int j0 = 0; j0 < n; j0 += kJChunk) {
    __syncthreads();
    for (int k = threadIdx.x; k < kJChunk * kLanes; k += blockDim.x) {
      const int jj = k / kLanes, l = k % kLanes;
      const int j = j0 + jj, vv = blockIdx.y * kLanes + l;
      double xv = 0.0;
      if (j < n && vv < nView) {
        xv = static_cast<double>(in[static_cast<size_t>(vv) * n + j]);
        if (w) xv *= w[j];
      }
      s_x[jj][l] = xv;
    }
    __syncthreads();
    const int jn = min(kJChunk, n - j0);
    int jj = 0;
    // 8 steps at a time: output o at step j + u takes tap i0 + o - j - u +
    // n - 1, so one window of kConvOut + 7 taps serves the block (T[7 + o -
    // u]); every sum still adds its products in input order
    for (; jj + 8 <= jn; jj += 8) {
      const int base = i0 - (j0 + jj) + n - 1 - 7 - ib;  // local tap index of T[0]
      double T[kConvOut + 7];
#pragma unroll
      for (int m = 0; m < kConvOut + 7; m++) T[m] = s_h[max(base + m, 0)];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const double xv = s_x[jj + u][lane];
#pragma unroll
        for (int o = 0; o < kConvOut; o++)
          if (i0 + o < n) acc[o] += xv * T[7 + o - u];
      }
    }
    for (; jj < jn; jj++) {
      const double xv = s_x[jj][lane];
      const int j = j0 + jj;
#pragma unroll
      for (int o = 0; o < kConvOut; o++) {
        const int i = i0 + o;
        if (i < n) acc[o] += xv * s_h[i - j + n - 1 - ib];
      }
    }
  }
  if (v < nView) {
#pragma unroll
    for (int o = 0; o < kConvOut; o++) {
      const int i = i0 + o;
      if (i < n) out[static_cast<size_t>(v) * outStride + i] = static_cast<TO>(acc[o] * dx);
    }
  }
}

// ----------------------------------------------------------- exact DFT
// Forward: Z[v][k] = (1/N) sum_{j<n} x[v][j] * (c[kj] + i s[kj]), then
// multiplied by filter[k]   (procsignal.cpp:570-576, 706-728, 778-801)
__global__ __launch_bounds__(kLanes * kWaves) void dft_fwd_kernel(
    const float* __restrict__ in, const double* __restrict__ w, const double* __restrict__ tc,
    const double* __restrict__ ts, const double* __restrict__ filt, double2* __restrict__ Z,
    int nView, int n, int N) {
  __shared__ double s_x[kJChunk][kLanes];
  const int lane = threadIdx.x & (kLanes - 1), wave = threadIdx.x / kLanes;
  const int v = blockIdx.y * kLanes + lane;
  const int k0 = (blockIdx.x * kWaves + wave) * kOutPerWave;
  double sr[kOutPerWave], si[kOutPerWave];
#pragma unroll
  for (int o = 0; o < kOutPerWave; o++) sr[o] = si[o] = 0.0;
  for (int j0 = 0; j0 < n; j0 += kJChunk) {
    __syncthreads();
    for (int k = threadIdx.x; k < kJChunk * kLanes; k += blockDim.x) {
      const int jj = k / kLanes, l = k % kLanes;
      const int j = j0 + jj, vv = blockIdx.y * kLanes + l;
      double xv = 0.0;
      if (j < n && vv < nView) {
        xv = static_cast<double>(in[static_cast<size_t>(vv) * n + j]);
        if (w) xv *= w[j];
      }
      s_x[jj][l] = xv;
    }
    __syncthreads();
    const int jn = min(kJChunk, n - j0);
    for (int jj = 0; jj < jn; jj++) {
      const double xv = s_x[jj][lane];
      const int j = j0 + jj;
#pragma unroll
      for (int o = 0; o < kOutPerWave; o++) {
        const int kk = k0 + o;
        if (kk < N) {
          const size_t t = static_cast<size_t>(kk) * j;
          sr[o] += xv * tc[t];
          si[o] += xv * ts[t];
        }
      }
    }
  }
  if (v < nView) {
#pragma unroll
    for (int o = 0; o < kOutPerWave; o++) {
      const int kk = k0 + o;
      if (kk < N) {
        double re = sr[o] / N, im = si[o] / N;
        re *= filt[kk];
        im *= filt[kk];
        Z[static_cast<size_t>(v) * N + kk] = make_double2(re, im);
      }
    }
  }
}

// Inverse: out[v][i] = sum_{k<N} Re(Z[k]) c[ik] - Im(Z[k]) s[ik], i < n
// (procsignal.cpp:577-580, 756-774, 843-864)
template <class TO>
__global__ __launch_bounds__(kLanes * kWaves) void dft_inv_kernel(
    const double2* __restrict__ Z, const double* __restrict__ tc, const double* __restrict__ ts,
    TO* __restrict__ out, int nView, int n, int N, int outStride) {
  __shared__ double2 s_z[kKChunk][kLanes];
  const int lane = threadIdx.x & (kLanes - 1), wave = threadIdx.x / kLanes;
  const int v = blockIdx.y * kLanes + lane;
  const int i0 = (blockIdx.x * kWaves + wave) * kOutPerWave;
  double acc[kOutPerWave];
#pragma unroll
  for (int o = 0; o < kOutPerWave; o++) acc[o] = 0.0;
  for (int k0 = 0; k0 < N; k0 += kKChunk) {
    __syncthreads();
    for (int k = threadIdx.x; k < kKChunk * kLanes; k += blockDim.x) {
      const int kk = k / kLanes, l = k % kLanes;
      const int vv = blockIdx.y * kLanes + l;
      s_z[kk][l] = (k0 + kk < N && vv < nView) ? Z[static_cast<size_t>(vv) * N + k0 + kk]
                                               : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int kn = min(kKChunk, N - k0);
    for (int kk = 0; kk < kn; kk++) {
      const double2 z = s_z[kk][lane];
      const int k = k0 + kk;
#pragma unroll
      for (int o = 0; o < kOutPerWave; o++) {
        const int i = i0 + o;
        if (i < n) {
          const size_t t = static_cast<size_t>(i) * k;
          acc[o] += z.x * tc[t] - z.y * ts[t];
        }
      }
    }
  }
  if (v < nView) {
#pragma unroll
    for (int o = 0; o < kOutPerWave; o++) {
      const int i = i0 + o;
      if (i < n) out[static_cast<size_t>(v) * outStride + i] = static_cast<TO>(acc[o]);
    }
  }
}

// ------------------------------------------------------------ LDS FFT
using dev::Cplx;

template <class T, class TO, int R>
__global__ __launch_bounds__(256) void fft_filter_kernel(
    const float* __restrict__ in, const double* __restrict__ w, const T* __restrict__ hre,
    const T* __restrict__ him, const Cplx<T>* __restrict__ twg, TO* __restrict__ out, int nView,
    int n, int N, int logN, int outStride) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  Cplx<T>* z = reinterpret_cast<Cplx<T>*>(smem_raw);
  Cplx<T>* tw = z + N;
  const int va = 2 * blockIdx.x, vb = va + 1 < nView ? va + 1 : -1;
  dev::fft_filter_pair<T, R>(in, w, hre, him, twg, va, vb, n, N, logN, z, tw);
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const Cplx<T> r = z[j];
    out[static_cast<size_t>(va) * outStride + j] = static_cast<TO>(r.x);
    if (vb >= 0) out[static_cast<size_t>(vb) * outStride + j] = static_cast<TO>(r.y);
  }
}


// FFT filtering of two views whose output is the interval coefficients of
// the backprojection's interpolation (mixed_device.h), formed straight from
// the filtered rows in LDS: the fused single-slice FBP writes no filtered
// rows.
template <int INTERP, bool LAST_INCL, int R>
__global__ __launch_bounds__(256) void fft_coef_kernel(
    const float* __restrict__ in, const double* __restrict__ w, const float* __restrict__ hre,
    const float* __restrict__ him, const Cplx<float>* __restrict__ twg, void* __restrict__ coef,
    int nView, int n, int N, int logN) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  Cplx<float>* z = reinterpret_cast<Cplx<float>*>(smem_raw);
  Cplx<float>* tw = z + N;
  const int va = 2 * blockIdx.x, vb = va + 1 < nView ? va + 1 : -1;
  dev::fft_filter_pair<float, R>(in, w, hre, him, twg, va, vb, n, N, logN, z, tw);
  using CT = dev::CoefT<INTERP>;
  CT* c = static_cast<CT*>(coef);
  auto ra = [z](int j) { return z[j].x; };
  auto rb = [z](int j) { return z[j].y; };
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    c[static_cast<size_t>(va) * n + j] = dev::interval_coef<INTERP, LAST_INCL>(ra, j, n);
    if (vb >= 0) c[static_cast<size_t>(vb) * n + j] = dev::interval_coef<INTERP, LAST_INCL>(rb, j, n);
  }
}

// The FP32 kernels with the transform length at compile time (64 ... 2048
// points): same results as fft_filter_kernel / fft_coef_kernel.
template <int LOGN>
__global__ __launch_bounds__(256) void fft_filter_kernel_c(
    const float* __restrict__ in, const double* __restrict__ w, const float* __restrict__ hre,
    const float* __restrict__ him, const Cplx<float>* __restrict__ twg, float* __restrict__ out,
    int nView, int n, int outStride) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  Cplx<float>* z = reinterpret_cast<Cplx<float>*>(smem_raw);
  Cplx<float>* tw = z + (1 << LOGN);
  const int va = 2 * blockIdx.x, vb = va + 1 < nView ? va + 1 : -1;
  dev::fft_filter_pair_c<LOGN>(in, w, hre, him, twg, va, vb, n, z, tw);
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const Cplx<float> r = z[dev::fft_slot(j)];
    out[static_cast<size_t>(va) * outStride + j] = r.x;
    if (vb >= 0) out[static_cast<size_t>(vb) * outStride + j] = r.y;
  }
}

template <int INTERP, bool LAST_INCL, int LOGN>
__global__ __launch_bounds__(256) void fft_coef_kernel_c(
    const float* __restrict__ in, const double* __restrict__ w, const float* __restrict__ hre,
    const float* __restrict__ him, const Cplx<float>* __restrict__ twg, void* __restrict__ coef,
    int nView, int n) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  Cplx<float>* z = reinterpret_cast<Cplx<float>*>(smem_raw);
  Cplx<float>* tw = z + (1 << LOGN);
  const int va = 2 * blockIdx.x, vb = va + 1 < nView ? va + 1 : -1;
  dev::fft_filter_pair_c<LOGN>(in, w, hre, him, twg, va, vb, n, z, tw);
  using CT = dev::CoefT<INTERP>;
  CT* c = static_cast<CT*>(coef);
  auto rz = [z](int j) { return z[dev::fft_slot(j)]; };
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    CT ca, cb;
    dev::interval_coef2<INTERP, LAST_INCL>(rz, j, n, ca, cb);
    dev::store_stream(&c[static_cast<size_t>(va) * n + j], ca);
    if (vb >= 0) dev::store_stream(&c[static_cast<size_t>(vb) * n + j], cb);
  }
}

// ------------------------------------------- Toeplitz GEMM on the MFMA units
// Mixed-precision convolution (procsignal.cpp:659-673 as a matrix product):
//   Y[v][i] = sum_j X[v][j] * T[j][i],  T[j][i] = dx * h[i - j + n - 1] = g[j - i + n - 1]
// M = views, N = outputs, K = inputs.  X and g are split into three BF16
// parts (hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid)) and the
// product keeps every term down to 2^-16 relative, hi*hi + hi*mid + mid*hi +
// hi*lo + lo*hi + mid*mid, with FP32 accumulation: six
// v_mfma_f32_32x32x16_bf16 per tile and k-step, a product accurate to about
// 2^-24 (a two-part split leaves 2^-17 per product, which the ramp filter's
// cancellation amplifies past 1e-4 in the image).
//
// A first kernel splits X once (after the preweight) into three BF16
// planes in fragment order, rows padded with zeros to a multiple of 128
// inputs.  In the product a workgroup (4 waves) computes a 128-view x
// 64-output tile; wave w owns rows w*32.. and two 32-column MFMA tiles.  The
// Toeplitz operand is never materialised: lane l of a B fragment needs the 8
// consecutive values g[q..q+8) with q = k0 - col + n - 1, so the plan keeps a
// record table rec[part][q + kRecOff] = g_part[q..q+8) (16 bytes, zero outside
// g) and every fragment is one 16-byte load that the L1/L2 serve.  Tile j
// (32 j columns further) at k-step s needs exactly tile 0's records of step
// s - 2 j, so each step loads records for tile 0 only, into a ring: six
// 16-byte loads per lane feed 12 MFMAs, at 3 waves per SIMD (four tiles per
// wave halve the loads per MFMA but leave 2 waves per SIMD, and run slower).
// Projection fragments are loaded three steps ahead (their first use comes
// from HBM), records one step ahead (they stay in L2); there is no LDS and
// no barrier.  The four waves of a workgroup read the same records (through
// L1) and distinct rows.  Workgroups are numbered so that each XCD runs a
// contiguous range of row blocks: the projection rows are read through one
// XCD's L2.
#pragma clang fp contract(fast)
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));

constexpr int kTM = 128;     // rows (views) per workgroup: 32 per wave
constexpr int kToeTJ = 2;    // 32-column MFMA tiles per wave (4: fewer loads, but 2 waves per SIMD)
constexpr int kRecOff = 192;  // record of g index q sits at q + kRecOff

int toeplitz_records(int n) { return 2 * n + 256; }
int toeplitz_pad(int n) { return (n + 127) / 128 * 128; }
size_t toeplitz_split_bytes(int n, int nView) {
  return 3 * static_cast<size_t>((nView + 31) / 32) * 32 * toeplitz_pad(n) * sizeof(unsigned short);
}

__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = static_cast<__bf16>(x);
  const float r1 = x - static_cast<float>(h);
  m = static_cast<__bf16>(r1);
  l = static_cast<__bf16>(r1 - static_cast<float>(m));
}

// The planes in MFMA fragment order: per part, per block of 32 rows, per
// k-step of 16 inputs, 1 KB holding lane l's 8 values (row l % 32, inputs
// 8 (l / 32) ..) at l * 16 bytes, so a fragment is one coalesced 1 KB load.
// Rows past the last view (up to the block's end) and inputs past n are
// zero.  A workgroup takes 32 rows x 64 inputs: coalesced row reads (times
// the preweight, PW), the split into LDS, then 4 KB of contiguous fragments
// per part.
template <bool PW>
__global__ __launch_bounds__(256) void toeplitz_split_kernel(const float* __restrict__ in,
                                                             const float* __restrict__ w32,
                                                             u16x8_t* __restrict__ planes, int nView, int n,
                                                             int nPad) {
  __shared__ __align__(16) unsigned short sp[3][32][64 + 8];  // [part][row][input] (padded rows)
  const int nKC = nPad / 64, nRB = (nView + 31) / 32;
  const int rb = blockIdx.x / nKC, kc = blockIdx.x - rb * nKC, t = threadIdx.x;
  const bool vec = (n & 3) == 0;
#pragma unroll
  for (int it = 0; it < 2; it++) {
    const int idx = it * 256 + t, r = idx >> 4, c = (idx & 15) * 4;
    const int v = rb * 32 + r, k = kc * 64 + c;
    float x[4] = {0.f, 0.f, 0.f, 0.f};
    if (v < nView) {
      const float* row = in + static_cast<size_t>(v) * n;
      if (vec && k + 4 <= n) {
        const float4 q = *reinterpret_cast<const float4*>(row + k);
        x[0] = q.x; x[1] = q.y; x[2] = q.z; x[3] = q.w;
      } else {
#pragma unroll
        for (int e = 0; e < 4; e++) x[e] = k + e < n ? row[k + e] : 0.f;
      }
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      if constexpr (PW)
        if (k + e < n) x[e] *= w32[k + e];
      __bf16 h, m, l;
      split3(x[e], h, m, l);
      sp[0][r][c + e] = __builtin_bit_cast(unsigned short, h);
      sp[1][r][c + e] = __builtin_bit_cast(unsigned short, m);
      sp[2][r][c + e] = __builtin_bit_cast(unsigned short, l);
    }
  }
  __syncthreads();
  // unit t: k-step kc * 4 + t / 64, lane t % 64 (row (t % 64) % 32, inputs 8 ((t % 64) / 32) ..)
  const int j = t >> 6, l = t & 63;
  const size_t plane = static_cast<size_t>(nRB) * (nPad / 16) * 64;  // 16-byte units
  const size_t at = (static_cast<size_t>(rb) * (nPad / 16) + kc * 4) * 64 + t;
#pragma unroll
  for (int part = 0; part < 3; part++)
    planes[part * plane + at] = *reinterpret_cast<const u16x8_t*>(&sp[part][l & 31][j * 16 + (l >> 5) * 8]);
}

// The six products of the four tiles (tile j takes records b[j]), term by
// term, smallest terms first: consecutive MFMAs accumulate into different
// tiles, so none waits for the one before it.
template <int TJ>
__device__ __forceinline__ void toe_products(floatx16_t* acc, const u16x8_t* a, const u16x8_t* const* b) {
  const bf16x8_t ah = __builtin_bit_cast(bf16x8_t, a[0]);
  const bf16x8_t am = __builtin_bit_cast(bf16x8_t, a[1]);
  const bf16x8_t al = __builtin_bit_cast(bf16x8_t, a[2]);
#define CTSIM_TOE_TERM(A, PART)                                                                     \
  _Pragma("unroll") for (int j = 0; j < TJ; j++) acc[j] =                                           \
      __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, __builtin_bit_cast(bf16x8_t, b[j][PART]), acc[j], 0, 0, 0);
  CTSIM_TOE_TERM(am, 1)
  CTSIM_TOE_TERM(al, 0)
  CTSIM_TOE_TERM(ah, 2)
  CTSIM_TOE_TERM(am, 0)
  CTSIM_TOE_TERM(ah, 1)
  CTSIM_TOE_TERM(ah, 0)
#undef CTSIM_TOE_TERM
}

// TJ 32-column tiles per wave; grid: 8 * ceil(tiles / 8) workgroups of 256
// lanes, tiles = ceil(n / (32 TJ)) * ceil(nView / 128); planes from
// toeplitz_split_kernel (nPad = toeplitz_pad(n)); rec: 3 * nrec records (hi,
// mid, lo), nrec = toeplitz_records(n).
template <int TJ>
__global__ __launch_bounds__(256) void toeplitz_mfma_kernel(const unsigned short* __restrict__ planes,
                                                            const u16x8_t* __restrict__ rec, int nrec,
                                                            float* __restrict__ out, int nView, int n, int nPad,
                                                            int outStride) {
  constexpr int TN = 32 * TJ, RS = 2 * TJ, UN = RS > 4 ? RS : 4;  // columns, record ring, unroll
  const int ncb = (n + TN - 1) / TN;
  const int tiles = ncb * ((nView + kTM - 1) / kTM);
  const int perXcd = (tiles + 7) / 8;
  const int t = (blockIdx.x & 7) * perXcd + (blockIdx.x >> 3);
  if (t >= tiles) return;
  const int m0 = (t / ncb) * kTM, n0 = (t % ncb) * TN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kg = lane >> 5;
  // one resource per plane: row blocks past the last read as zero
  const int nRB = (nView + 31) / 32, nSteps = nPad / 16;  // nSteps: a multiple of 8
  const size_t planeBytes = static_cast<size_t>(nRB) * nSteps * 1024;
  __amdgpu_buffer_rsrc_t rsA[3];
#pragma unroll
  for (int part = 0; part < 3; part++)
    rsA[part] = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<unsigned char*>(reinterpret_cast<const unsigned char*>(planes) + part * planeBytes), 0,
        static_cast<int>(planeBytes), 0x00020000);
  const __amdgpu_buffer_rsrc_t rsB = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<u16x8_t*>(rec), 0, 3 * nrec * 16, 0x00020000);
  const int aOff = (m0 / 32 + wave) * nSteps * 1024 + lane * 16;  // bytes
  const int bIdx = 8 * kg - (n0 + (lane & 31)) + n - 1 + kRecOff;
  auto loadA = [&](u16x8_t* a, int st) {
#pragma unroll
    for (int part = 0; part < 3; part++)
      a[part] = __builtin_bit_cast(u16x8_t, __builtin_amdgcn_raw_buffer_load_b128(rsA[part], aOff + st * 1024, 0, 0));
  };
  auto loadB = [&](u16x8_t* b, int st) {  // past the table's ends: zeros
#pragma unroll
    for (int part = 0; part < 3; part++)
      b[part] = __builtin_bit_cast(
          u16x8_t, __builtin_amdgcn_raw_buffer_load_b128(rsB, (part * nrec + bIdx + 16 * st) * 16, 0, 0));
  };
  floatx16_t acc[TJ] = {};
  // A: steps s .. s+3 at [s & 3] (the projection rows come from HBM on
  // first use); R: tile-0 records of steps s-2(TJ-1) .. s+1 at [s % RS]
  // (tile j at step s is tile 0 at step s - 2j; the records stay in L2)
  u16x8_t A[4][3], R[RS][3];
#pragma unroll
  for (int st = 2 - RS; st <= 0; st++) loadB(R[st & (RS - 1)], st);
#pragma unroll
  for (int st = 0; st < 3; st++) loadA(A[st], st);
  for (int s0 = 0; s0 < nSteps; s0 += UN) {
#pragma unroll
    for (int u = 0; u < UN; u++) {
      const int st = s0 + u;
      // records first: the wait for them (loads complete in issue order)
      // then leaves the projection fragments in flight
      loadB(R[(u + 1) & (RS - 1)], st + 1);
      loadA(A[(u + 3) & 3], st + 3);  // past the last step: zeros or unused
      __builtin_amdgcn_sched_barrier(0);  // keep the next step's loads ahead of this step's MFMAs
      const u16x8_t* bj[TJ];
#pragma unroll
      for (int j = 0; j < TJ; j++) bj[j] = R[(u - 2 * j) & (RS - 1)];
      toe_products<TJ>(acc, A[u & 3], bj);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // D layout of the 32x32 tile: register r of lane l is row 8(r/4) + 4(l/32) + r%4,
  // column l%32.
#pragma unroll
  for (int j = 0; j < TJ; j++) {
    const int col = n0 + j * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int rr = m0 + wave * 32 + 8 * (r / 4) + 4 * kg + (r % 4);
      if (rr < nView && col < n) out[static_cast<size_t>(rr) * outStride + col] = acc[j][r];
    }
  }
}
#pragma clang fp contract(off)

// ------------------------------------------------------------ launchers

template <class T, class TO>
int launch_fft(ctsim_ctx* ctx, const FilterPlan& p, const float* d_in, int nView, TO* d_out,
               int outStride) {
  const int N = p.nfp;
  int logN = 0;
  while ((1 << logN) < N) logN++;
  if ((1 << logN) != N)
    return fail(CTSIM_ERR_UNSUPPORTED, "FFT filtering needs a power-of-two transform length "
                                       "(zeropad >= 1)");
  const size_t lds = static_cast<size_t>(N + dev::fft_tw_entries<T>(N)) * sizeof(Cplx<T>);
  if (lds > kLdsPerGroup) return fail(CTSIM_ERR_UNSUPPORTED, "FFT filtering: transform too long for LDS");
  const T* hre = (sizeof(T) == 8 ? p.d_hmul64 : p.d_hmul32).template as<T>();
  const T* him = hre + (N / 2 + 1);
  const Cplx<T>* tw = (sizeof(T) == 8 ? p.d_tw64 : p.d_tw32).template as<Cplx<T>>();
  const double* w = p.preweight.empty() ? nullptr : p.d_weight.as<double>();
  const int grid = (nView + 1) / 2;
  const int R = (N / 2 + 255) / 256;
  if constexpr (sizeof(T) == 4 && sizeof(TO) == 4) {
    if (dev::fft_c_length(logN)) {
      const size_t ldsC = dev::fft_c_lds(N);
#define CTSIM_FFT_C(L)                                                                         \
  case L:                                                                                      \
    hipLaunchKernelGGL((fft_filter_kernel_c<L>), dim3(grid), dim3(256), ldsC, ctx->stream, d_in, w, \
                       hre, him, tw, reinterpret_cast<float*>(d_out), nView, p.n, outStride);  \
    break;
      switch (logN) {
        CTSIM_FFT_C(6)
        CTSIM_FFT_C(7)
        CTSIM_FFT_C(8)
        CTSIM_FFT_C(9)
        CTSIM_FFT_C(10)
        CTSIM_FFT_C(11)
        CTSIM_FFT_C(12)
      }
#undef CTSIM_FFT_C
      CTSIM_HIP(hipGetLastError());
      return CTSIM_OK;
    }
  }
#define CTSIM_FFT_CASE(RR)                                                                     \
  case RR:                                                                                     \
    hipLaunchKernelGGL((fft_filter_kernel<T, TO, RR>), dim3(grid), dim3(256), lds, ctx->stream, \
                       d_in, w, hre, him, tw, d_out, nView, p.n, N, logN, outStride);          \
    break;
  switch (R) {
    CTSIM_FFT_CASE(1)
    CTSIM_FFT_CASE(2)
    CTSIM_FFT_CASE(4)
    CTSIM_FFT_CASE(8)
    CTSIM_FFT_CASE(16)
    default: return fail(CTSIM_ERR_UNSUPPORTED, "FFT filtering: unsupported transform length");
  }
#undef CTSIM_FFT_CASE
  CTSIM_HIP(hipGetLastError());
  return CTSIM_OK;
}

}  // namespace

int filter_upload(FilterPlan& p) {
  CTSIM_CHECK(p.d_filter.reserve(p.filter.size() * sizeof(double)));
  CTSIM_HIP(hipMemcpy(p.d_filter.ptr, p.filter.data(), p.filter.size() * sizeof(double),
                      hipMemcpyHostToDevice));
  if (!p.preweight.empty()) {
    CTSIM_CHECK(p.d_weight.reserve(p.preweight.size() * sizeof(double)));
    CTSIM_HIP(hipMemcpy(p.d_weight.ptr, p.preweight.data(), p.preweight.size() * sizeof(double),
                        hipMemcpyHostToDevice));
    std::vector<float> w32(p.preweight.begin(), p.preweight.end());
    CTSIM_CHECK(p.d_weight32.reserve(w32.size() * sizeof(float)));
    CTSIM_HIP(hipMemcpy(p.d_weight32.ptr, w32.data(), w32.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  if (p.method == METHOD_CONVOLUTION) {
    // Toeplitz operand of the mixed convolution: g[q] = dx * h[2n-2-q],
    // split into bf16 hi, mid, lo from its double value, stored as records of
    // 8 consecutive values: record i of a part holds g_part[i - kRecOff + 0..7]
    const int nq = 2 * p.n - 1, nrec = toeplitz_records(p.n);
    std::vector<unsigned short> g16(3 * static_cast<size_t>(nq));
    for (int q = 0; q < nq; q++) {
      const double g = p.inc * p.filter[2 * p.n - 2 - q];
      const __bf16 hi = static_cast<__bf16>(static_cast<float>(g));
      const double r1 = g - static_cast<double>(static_cast<float>(hi));
      const __bf16 mid = static_cast<__bf16>(static_cast<float>(r1));
      const __bf16 lo = static_cast<__bf16>(static_cast<float>(r1 - static_cast<double>(static_cast<float>(mid))));
      g16[q] = __builtin_bit_cast(unsigned short, hi);
      g16[nq + q] = __builtin_bit_cast(unsigned short, mid);
      g16[2 * static_cast<size_t>(nq) + q] = __builtin_bit_cast(unsigned short, lo);
    }
    std::vector<unsigned short> recs(3 * static_cast<size_t>(nrec) * 8, 0);
    for (int part = 0; part < 3; part++)
      for (int i = 0; i < nrec; i++)
        for (int e = 0; e < 8; e++) {
          const int q = i - kRecOff + e;
          if (q >= 0 && q < nq)
            recs[(static_cast<size_t>(part) * nrec + i) * 8 + e] = g16[static_cast<size_t>(part) * nq + q];
        }
    CTSIM_CHECK(p.d_toeplitz.reserve(recs.size() * sizeof(unsigned short)));
    CTSIM_HIP(hipMemcpy(p.d_toeplitz.ptr, recs.data(), recs.size() * sizeof(unsigned short),
                        hipMemcpyHostToDevice));
  }
  if (!p.hre.empty()) {
    const int N = p.nfp, H = N / 2 + 1;
    std::vector<double> m64(2 * H);
    std::vector<float> m32(2 * H);
    for (int k = 0; k < H; k++) {
      m64[k] = p.hre[k];
      m64[H + k] = p.him[k];
      m32[k] = static_cast<float>(p.hre[k]);
      m32[H + k] = static_cast<float>(p.him[k]);
    }
    CTSIM_CHECK(p.d_hmul64.reserve(m64.size() * sizeof(double)));
    CTSIM_CHECK(p.d_hmul32.reserve(m32.size() * sizeof(float)));
    CTSIM_HIP(hipMemcpy(p.d_hmul64.ptr, m64.data(), m64.size() * sizeof(double), hipMemcpyHostToDevice));
    CTSIM_HIP(hipMemcpy(p.d_hmul32.ptr, m32.data(), m32.size() * sizeof(float), hipMemcpyHostToDevice));
    // exp(-2 pi i k / N) from FP64 host trig
    std::vector<double> t64(N);
    std::vector<float> t32(N);
    for (int k = 0; k < N / 2; k++) {
      const double a = -2.0 * kPi * k / N;
      t64[2 * k] = cos(a);
      t64[2 * k + 1] = sin(a);
      t32[2 * k] = static_cast<float>(t64[2 * k]);
      t32[2 * k + 1] = static_cast<float>(t64[2 * k + 1]);
    }
    dev::fft_append_pass_twiddles(t32, N);
    CTSIM_CHECK(p.d_tw64.reserve(t64.size() * sizeof(double)));
    CTSIM_CHECK(p.d_tw32.reserve(t32.size() * sizeof(float)));
    CTSIM_HIP(hipMemcpy(p.d_tw64.ptr, t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice));
    CTSIM_HIP(hipMemcpy(p.d_tw32.ptr, t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  return CTSIM_OK;
}

static int upload_dft_tables(FilterPlan& p) {
  if (p.d_dft.ptr) return CTSIM_OK;
  filter_build_dft_tables(p);
  const size_t nk = p.tfc.size();
  CTSIM_CHECK(p.d_dft.reserve(4 * nk * sizeof(double)));
  double* d = p.d_dft.as<double>();
  CTSIM_HIP(hipMemcpy(d, p.tfc.data(), nk * sizeof(double), hipMemcpyHostToDevice));
  CTSIM_HIP(hipMemcpy(d + nk, p.tfs.data(), nk * sizeof(double), hipMemcpyHostToDevice));
  CTSIM_HIP(hipMemcpy(d + 2 * nk, p.tbc.data(), nk * sizeof(double), hipMemcpyHostToDevice));
  CTSIM_HIP(hipMemcpy(d + 3 * nk, p.tbs.data(), nk * sizeof(double), hipMemcpyHostToDevice));
  // the host copies are no longer needed
  std::vector<double>().swap(p.tfc);
  std::vector<double>().swap(p.tfs);
  std::vector<double>().swap(p.tbc);
  std::vector<double>().swap(p.tbs);
  return CTSIM_OK;
}

bool filter_can_coef(const FilterPlan& p) {
  const int N = p.nfp;
  return p.method != METHOD_CONVOLUTION && !p.hre.empty() && p.preinterp == 1 && N >= 2 && N <= 4096 &&
         (N & (N - 1)) == 0;
}

int filter_launch_coef(ctsim_ctx* ctx, FilterPlan& p, const float* d_in, int nView, void* d_coef,
                       int interp, bool lastIncl) {
  if (nView <= 0) return CTSIM_OK;
  if (!filter_can_coef(p))
    return fail(CTSIM_ERR_UNSUPPORTED, "coefficient output needs a power-of-two FFT filter up to 4096 points");
  const int N = p.nfp;
  int logN = 0;
  while ((1 << logN) < N) logN++;
  const size_t lds = static_cast<size_t>(N + dev::fft_tw_entries<float>(N)) * sizeof(Cplx<float>);
  const float* hre = p.d_hmul32.as<float>();
  const float* him = hre + (N / 2 + 1);
  const Cplx<float>* tw = p.d_tw32.as<Cplx<float>>();
  const double* w = p.preweight.empty() ? nullptr : p.d_weight.as<double>();
  const int grid = (nView + 1) / 2;
  const int R = (N / 2 + 255) / 256;
  if (dev::fft_c_length(logN)) {
    auto goc = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(grid), dim3(256), dev::fft_c_lds(N), ctx->stream, d_in, w, hre, him,
                         tw, d_coef, nView, p.n);
    };
#define CTSIM_COEF_C(L)                                                                       \
  case L:                                                                                     \
    if (interp == INTERP_CUBIC) goc(fft_coef_kernel_c<INTERP_CUBIC, false, L>);              \
    else if (lastIncl) goc(fft_coef_kernel_c<INTERP_LINEAR, true, L>);                       \
    else goc(fft_coef_kernel_c<INTERP_LINEAR, false, L>);                                     \
    break;
    switch (logN) {
      CTSIM_COEF_C(6)
      CTSIM_COEF_C(7)
      CTSIM_COEF_C(8)
      CTSIM_COEF_C(9)
      CTSIM_COEF_C(10)
      CTSIM_COEF_C(11)
      CTSIM_COEF_C(12)
    }
#undef CTSIM_COEF_C
    CTSIM_HIP(hipGetLastError());
    return CTSIM_OK;
  }
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, ctx->stream, d_in, w, hre, him, tw, d_coef,
                       nView, p.n, N, logN);
  };
#define CTSIM_COEF_CASE(RR)                                                                   \
  case RR:                                                                                    \
    if (interp == INTERP_CUBIC) go(fft_coef_kernel<INTERP_CUBIC, false, RR>);                 \
    else if (lastIncl) go(fft_coef_kernel<INTERP_LINEAR, true, RR>);                          \
    else go(fft_coef_kernel<INTERP_LINEAR, false, RR>);                                        \
    break;
  switch (R) {
    CTSIM_COEF_CASE(1)
    CTSIM_COEF_CASE(2)
    CTSIM_COEF_CASE(4)
    CTSIM_COEF_CASE(8)
    default: return fail(CTSIM_ERR_UNSUPPORTED, "FFT filtering: unsupported transform length");
  }
#undef CTSIM_COEF_CASE
  CTSIM_HIP(hipGetLastError());
  return CTSIM_OK;
}

size_t filter_work_bytes(const FilterPlan& p, int precision, int nView) {
  if (p.method == METHOD_CONVOLUTION && precision != CTSIM_PREC_F64) return toeplitz_split_bytes(p.n, nView);
  return 0;
}

int filter_launch(ctsim_ctx* ctx, FilterPlan& p, int precision, const float* d_in, int nView,
                  void* d_out, int outStride, void* d_work) {
  if (nView <= 0) return CTSIM_OK;
  const hipStream_t s = ctx->stream;
  const double* w = p.preweight.empty() ? nullptr : p.d_weight.as<double>();
  const bool f64 = precision == CTSIM_PREC_F64;
  if (p.method == METHOD_CONVOLUTION && !f64) {
    const int tiles = ((p.n + 32 * kToeTJ - 1) / (32 * kToeTJ)) * ((nView + kTM - 1) / kTM);
    const int nPad = toeplitz_pad(p.n);
    if (toeplitz_split_bytes(p.n, nView) / 3 > INT32_MAX - 4096 || 2LL * p.n + 512 > INT32_MAX / 48)
      return fail(CTSIM_ERR_UNSUPPORTED, "mixed convolution: detector count too large");
    if (!d_work) {
      CTSIM_CHECK(ctx->scratch.reserve(toeplitz_split_bytes(p.n, nView)));
      d_work = ctx->scratch.ptr;
    }
    u16x8_t* planes = static_cast<u16x8_t*>(d_work);
    const float* w32 = p.preweight.empty() ? nullptr : p.d_weight32.as<float>();
    const dim3 gs(static_cast<unsigned>((nView + 31) / 32 * (nPad / 64)));
    if (w32)
      hipLaunchKernelGGL((toeplitz_split_kernel<true>), gs, dim3(256), 0, s, d_in, w32, planes, nView, p.n, nPad);
    else
      hipLaunchKernelGGL((toeplitz_split_kernel<false>), gs, dim3(256), 0, s, d_in, w32, planes, nView, p.n, nPad);
    hipLaunchKernelGGL(toeplitz_mfma_kernel<kToeTJ>, dim3(8 * ((tiles + 7) / 8)), dim3(256), 0, s,
                       reinterpret_cast<const unsigned short*>(planes), p.d_toeplitz.as<u16x8_t>(),
                       toeplitz_records(p.n), static_cast<float*>(d_out), nView, p.n, nPad, outStride);
    CTSIM_HIP(hipGetLastError());
    return CTSIM_OK;
  }
  if (p.method == METHOD_CONVOLUTION) {
    const dim3 grid((p.n + kWaves * kConvOut - 1) / (kWaves * kConvOut), (nView + kLanes - 1) / kLanes);
    const size_t lds = static_cast<size_t>(p.n + kWaves * kConvOut - 1) * sizeof(double);
    if (lds + sizeof(double) * kJChunk * kLanes > 64 * 1024)
      return fail(CTSIM_ERR_UNSUPPORTED, "convolution filter: kernel too long for LDS");
    if (f64)
      hipLaunchKernelGGL((conv_seq_kernel<double>), grid, dim3(kLanes * kWaves), lds, s, d_in, w,
                         p.d_filter.as<double>(), static_cast<double*>(d_out), nView, p.n, p.inc, outStride);
    else
      hipLaunchKernelGGL((conv_seq_kernel<float>), grid, dim3(kLanes * kWaves), lds, s, d_in, w,
                         p.d_filter.as<double>(), static_cast<float*>(d_out), nView, p.n, p.inc, outStride);
    CTSIM_HIP(hipGetLastError());
    return CTSIM_OK;
  }
  if (filter_is_fourier(p) && f64) {
    const int N = p.nfp;
    if (N > 4096)
      return fail(CTSIM_ERR_UNSUPPORTED, "exact DFT filtering supports transforms up to 4096 points");
    CTSIM_CHECK(upload_dft_tables(p));
    const size_t nk = static_cast<size_t>(N - 1) * (N - 1) + 1;
    const double* tab = p.d_dft.as<double>();
    CTSIM_CHECK(ctx->scratch.reserve(static_cast<size_t>(nView) * N * sizeof(double2)));
    double2* Z = ctx->scratch.as<double2>();
    const int per = kWaves * kOutPerWave;
    hipLaunchKernelGGL(dft_fwd_kernel, dim3((N + per - 1) / per, (nView + kLanes - 1) / kLanes),
                       dim3(kLanes * kWaves), 0, s, d_in, w, tab, tab + nk,
                       p.d_filter.as<double>(), Z, nView, p.n, N);
    CTSIM_HIP(hipGetLastError());
    hipLaunchKernelGGL((dft_inv_kernel<double>), dim3((p.n + per - 1) / per, (nView + kLanes - 1) / kLanes),
                       dim3(kLanes * kWaves), 0, s, Z, tab + 2 * nk, tab + 3 * nk,
                       static_cast<double*>(d_out), nView, p.n, N, outStride);
    CTSIM_HIP(hipGetLastError());
    return CTSIM_OK;
  }
  if (f64) return launch_fft<double, double>(ctx, p, d_in, nView, static_cast<double*>(d_out), outStride);
  return launch_fft<float, float>(ctx, p, d_in, nView, static_cast<float*>(d_out), outStride);
}

}  // namespace ctsim

using namespace ctsim;

extern "C" {

int ctsim_procsignal_create(ctsim_ctx* ctx, ctsim_procsignal** out, const char* filterName,
                            const char* methodName, double bw, double signalIncrement, int n,
                            double param, const char* domainName, const char* generationName,
                            int zeropad, int preinterpolationFactor, int geometry,
                            double focalLength, double sourceDetectorLength) {
  if (!ctx || !out) return fail(CTSIM_ERR_INVALID, "ctsim_procsignal_create: NULL argument");
  *out = nullptr;
  CTSIM_CHECK(ctx->activate());
  auto* ps = new ctsim_procsignal();
  int rc = filter_plan_init(ps->plan, filterName, methodName, bw, signalIncrement, n, param,
                            domainName, generationName, zeropad, preinterpolationFactor, geometry,
                            focalLength, sourceDetectorLength);
  if (rc == CTSIM_OK) rc = filter_upload(ps->plan);
  if (rc != CTSIM_OK) {
    delete ps;
    return rc;
  }
  *out = ps;
  return CTSIM_OK;
}

void ctsim_procsignal_destroy(ctsim_procsignal* ps) { delete ps; }

int ctsim_procsignal_info(const ctsim_procsignal* ps, int* nFilterPoints, double* minMaxInc) {
  if (!ps) return fail(CTSIM_ERR_INVALID, "ctsim_procsignal_info: NULL argument");
  if (nFilterPoints) *nFilterPoints = ps->plan.nfp;
  if (minMaxInc) {
    minMaxInc[0] = ps->plan.fmin;
    minMaxInc[1] = ps->plan.fmax;
    minMaxInc[2] = ps->plan.finc;
  }
  return CTSIM_OK;
}

int ctsim_procsignal_get_filter(const ctsim_procsignal* ps, double* h_out, int capacity) {
  if (!ps || !h_out) return fail(CTSIM_ERR_INVALID, "ctsim_procsignal_get_filter: NULL argument");
  const int n = static_cast<int>(ps->plan.filter.size());
  for (int i = 0; i < n && i < capacity; i++) h_out[i] = ps->plan.filter[i];
  return n;
}

int ctsim_filter_views(ctsim_ctx* ctx, ctsim_procsignal* ps, int precision, const float* d_in,
                       int nView, void* d_out) {
  if (!ctx || !ps || (nView > 0 && (!d_in || !d_out)))
    return fail(CTSIM_ERR_INVALID, "ctsim_filter_views: NULL argument");
  if (precision != CTSIM_PREC_F64 && precision != CTSIM_PREC_MIXED)
    return fail(CTSIM_ERR_INVALID, "ctsim_filter_views: unknown precision");
  if (nView < 0) return fail(CTSIM_ERR_INVALID, "ctsim_filter_views: nView < 0");
  CTSIM_CHECK(ctx->activate());
  return filter_launch(ctx, ps->plan, precision, d_in, nView, d_out, ps->plan.n * ps->plan.preinterp);
}

}  // extern "C"
