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
constexpr int kKChunk = 32;       // spectrum entries staged per step (inverse DFT)

// ---------------------------------------------------- exact convolution
// out[v][i] = dx * sum_{j=0}^{n-1} x[v][j] * h[i - j + n - 1]  (procsignal.cpp:659-673)
template <class TO>
__global__ __launch_bounds__(kLanes * kWaves) void conv_seq_kernel(
    const float* __restrict__ in, const double* __restrict__ w, const double* __restrict__ h,
    TO* __restrict__ out, int nView, int n, double dx, int outStride) {
  __shared__ double s_x[kJChunk][kLanes];
  extern __shared__ double s_h[];  // 2n-1 taps
  const int lane = threadIdx.x & (kLanes - 1), wave = threadIdx.x / kLanes;
  const int v = blockIdx.y * kLanes + lane;
  const int i0 = (blockIdx.x * kWaves + wave) * kOutPerWave;
  for (int k = threadIdx.x; k < 2 * n - 1; k += blockDim.x) s_h[k] = h[k];
  double acc[kOutPerWave];
#pragma unroll
  for (int o = 0; o < kOutPerWave; o++) acc[o] = 0.0;
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
        const int i = i0 + o;
        const int hk = i - j + n - 1;
        if (i < n) acc[o] += xv * s_h[hk];
      }
    }
  }
  if (v < nView) {
#pragma unroll
    for (int o = 0; o < kOutPerWave; o++) {
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
  extern __shared__ unsigned char smem_raw[];
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
  extern __shared__ unsigned char smem_raw[];
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
  extern __shared__ unsigned char smem_raw[];
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
  extern __shared__ unsigned char smem_raw[];
  Cplx<float>* z = reinterpret_cast<Cplx<float>*>(smem_raw);
  Cplx<float>* tw = z + (1 << LOGN);
  const int va = 2 * blockIdx.x, vb = va + 1 < nView ? va + 1 : -1;
  dev::fft_filter_pair_c<LOGN>(in, w, hre, him, twg, va, vb, n, z, tw);
  using CT = dev::CoefT<INTERP>;
  CT* c = static_cast<CT*>(coef);
  auto ra = [z](int j) { return z[dev::fft_slot(j)].x; };
  auto rb = [z](int j) { return z[dev::fft_slot(j)].y; };
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    dev::store_stream(&c[static_cast<size_t>(va) * n + j], dev::interval_coef<INTERP, LAST_INCL>(ra, j, n));
    if (vb >= 0)
      dev::store_stream(&c[static_cast<size_t>(vb) * n + j], dev::interval_coef<INTERP, LAST_INCL>(rb, j, n));
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
// A workgroup (4 waves) computes a 64-view x 128-output tile; wave (wm, wn)
// owns rows wm*32.. and two 32-column MFMA tiles at wn*64.  The Toeplitz
// operand is never materialised: lane l of a B fragment needs the 8
// consecutive values g[q..q+8) with q = k0 - col + n - 1, so the plan keeps a
// record table rec[part][q + kRecOff] = g_part[q..q+8) (16 bytes, zero outside
// g) and every fragment is one 16-byte load that the L1/L2 serve.  Tile 1
// (32 columns further) of k-chunk j0 needs exactly tile 0's records of chunk
// j0 - 32, so each chunk loads records for tile 0 only and tile 1 reuses the
// previous chunk's.  Projection values and records of the next chunk are
// loaded while the current chunk's MFMAs run; there is no LDS and no barrier.
// Workgroups are numbered so that each XCD runs a contiguous range of row
// blocks: the projection rows are read through one XCD's L2.
#pragma clang fp contract(fast)
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));

constexpr int kTM = 64, kTN = 128, kTK = 32;
constexpr int kRecOff = 192;  // record of g index q sits at q + kRecOff

int toeplitz_records(int n) { return 2 * n + 256; }

struct ToeChunk {
  floatx4_t x[2][2];   // [k-step][half]: X[row][j0 + 16 s + 8 kg + 4 h + 0..3]
  floatx4_t w[2][2];   // the same positions of the preweight (PW only)
  u16x8_t b[2][3];     // [k-step][part]: tile-0 records
};

template <bool PW>
__device__ __forceinline__ void toe_load(ToeChunk& c, __amdgpu_buffer_rsrc_t rsA,
                                         __amdgpu_buffer_rsrc_t rsW, __amdgpu_buffer_rsrc_t rsB,
                                         int aOff, int wOff, int bIdx, int j0, int nrec) {
#pragma unroll
  for (int st = 0; st < 2; st++) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      c.x[st][h] = __builtin_bit_cast(
          floatx4_t, __builtin_amdgcn_raw_buffer_load_b128(rsA, (aOff + j0 + 16 * st + 4 * h) * 4, 0, 0));
      if constexpr (PW)
        c.w[st][h] = __builtin_bit_cast(
            floatx4_t, __builtin_amdgcn_raw_buffer_load_b128(rsW, (wOff + j0 + 16 * st + 4 * h) * 4, 0, 0));
    }
#pragma unroll
    for (int part = 0; part < 3; part++)
      c.b[st][part] = __builtin_bit_cast(
          u16x8_t, __builtin_amdgcn_raw_buffer_load_b128(rsB, (part * nrec + bIdx + j0 + 16 * st) * 16, 0, 0));
  }
}

__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = static_cast<__bf16>(x);
  const float r1 = x - static_cast<float>(h);
  m = static_cast<__bf16>(r1);
  l = static_cast<__bf16>(r1 - static_cast<float>(m));
}

__device__ __forceinline__ void toe_products(floatx16_t& acc, const bf16x8_t& ah, const bf16x8_t& am,
                                             const bf16x8_t& al, const u16x8_t* b) {
  const bf16x8_t bh = __builtin_bit_cast(bf16x8_t, b[0]);
  const bf16x8_t bm = __builtin_bit_cast(bf16x8_t, b[1]);
  const bf16x8_t bl = __builtin_bit_cast(bf16x8_t, b[2]);
  // smallest terms first
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bm, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bm, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
}

// grid: 8 * ceil(tiles / 8) workgroups of 256 lanes, tiles = ceil(n / 128) *
// ceil(nView / 64); rec: 3 * nrec records (hi, mid, lo), nrec = toeplitz_records(n);
// PW: the inputs are multiplied by the preweight w32 first.
template <bool PW>
__global__ __launch_bounds__(256) void toeplitz_mfma_kernel(
    const float* __restrict__ in, const float* __restrict__ w32, const u16x8_t* __restrict__ rec,
    int nrec, float* __restrict__ out, int nView, int n, int outStride) {
  const int ncb = (n + kTN - 1) / kTN;
  const int tiles = ncb * ((nView + kTM - 1) / kTM);
  const int perXcd = (tiles + 7) / 8;
  const int t = (blockIdx.x & 7) * perXcd + (blockIdx.x >> 3);
  if (t >= tiles) return;
  const int m0 = (t / ncb) * kTM, n0 = (t % ncb) * kTN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int kg = lane >> 5;
  const int rows = min(kTM, nView - m0);
  // rows past the last view read as zero (buffer range); values past the
  // row end (k >= n) are cleared in the last chunk
  const __amdgpu_buffer_rsrc_t rsA = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float*>(in + static_cast<size_t>(m0) * n), 0, rows * n * 4, 0x00020000);
  const __amdgpu_buffer_rsrc_t rsB = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<u16x8_t*>(rec), 0, 3 * nrec * 16, 0x00020000);
  const __amdgpu_buffer_rsrc_t rsW = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float*>(w32), 0, PW ? n * 4 : 0, 0x00020000);
  const int wOff = 8 * kg;
  const int aOff = (wm * 32 + (lane & 31)) * n + 8 * kg;
  const int bIdx = 8 * kg - (n0 + wn * 64 + (lane & 31)) + n - 1 + kRecOff;
  floatx16_t acc0 = {}, acc1 = {};
  ToeChunk cur, nxt;
  u16x8_t prev[2][3];  // tile-0 records of the previous chunk = tile 1's of this one
  toe_load<PW>(nxt, rsA, rsW, rsB, aOff, wOff, bIdx, -kTK, nrec);  // only its records are used
#pragma unroll
  for (int st = 0; st < 2; st++)
#pragma unroll
    for (int part = 0; part < 3; part++) prev[st][part] = nxt.b[st][part];
  toe_load<PW>(cur, rsA, rsW, rsB, aOff, wOff, bIdx, 0, nrec);
  const int nc = (n + kTK - 1) / kTK;
  for (int c = 0; c < nc; c++) {
    const int j0 = c * kTK;
    toe_load<PW>(nxt, rsA, rsW, rsB, aOff, wOff, bIdx, j0 + kTK, nrec);  // past the end: zeros, unused
    __builtin_amdgcn_sched_barrier(0);  // keep the next chunk's loads ahead of this chunk's MFMAs
#pragma unroll
    for (int st = 0; st < 2; st++) {
      const int k0 = j0 + 16 * st + 8 * kg;
      bf16x8_t ah, am, al;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        float x = cur.x[st][e >> 2][e & 3];
        if constexpr (PW) x *= cur.w[st][e >> 2][e & 3];
        if (j0 + kTK > n && k0 + e >= n) x = 0.f;
        __bf16 h, m, l;
        split3(x, h, m, l);
        ah[e] = h;
        am[e] = m;
        al[e] = l;
      }
      toe_products(acc0, ah, am, al, cur.b[st]);
      toe_products(acc1, ah, am, al, prev[st]);
    }
    __builtin_amdgcn_sched_barrier(0);  // rotate only after every MFMA is issued
#pragma unroll
    for (int st = 0; st < 2; st++)
#pragma unroll
      for (int part = 0; part < 3; part++) prev[st][part] = cur.b[st][part];
    cur = nxt;
  }
  // D layout of the 32x32 tile: register r of lane l is row 8(r/4) + 4(l/32) + r%4,
  // column l%32.
#pragma unroll
  for (int tile = 0; tile < 2; tile++) {
    const floatx16_t& acc = tile == 0 ? acc0 : acc1;
    const int col = n0 + wn * 64 + tile * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int rr = m0 + wm * 32 + 8 * (r / 4) + 4 * kg + (r % 4);
      if (rr < nView && col < n) out[static_cast<size_t>(rr) * outStride + col] = acc[r];
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

int filter_launch(ctsim_ctx* ctx, FilterPlan& p, int precision, const float* d_in, int nView,
                  void* d_out, int outStride) {
  if (nView <= 0) return CTSIM_OK;
  const hipStream_t s = ctx->stream;
  const double* w = p.preweight.empty() ? nullptr : p.d_weight.as<double>();
  const bool f64 = precision == CTSIM_PREC_F64;
  if (p.method == METHOD_CONVOLUTION && !f64) {
    const int tiles = ((p.n + kTN - 1) / kTN) * ((nView + kTM - 1) / kTM);
    if (static_cast<long long>(kTM) * p.n * 4 > INT32_MAX)
      return fail(CTSIM_ERR_UNSUPPORTED, "mixed convolution: detector count too large");
    const float* w32 = p.preweight.empty() ? nullptr : p.d_weight32.as<float>();
    auto kern = w32 ? toeplitz_mfma_kernel<true> : toeplitz_mfma_kernel<false>;
    hipLaunchKernelGGL(kern, dim3(8 * ((tiles + 7) / 8)), dim3(256), 0, s, d_in, w32,
                       p.d_toeplitz.as<u16x8_t>(), toeplitz_records(p.n),
                       static_cast<float*>(d_out), nView, p.n, outStride);
    CTSIM_HIP(hipGetLastError());
    return CTSIM_OK;
  }
  if (p.method == METHOD_CONVOLUTION) {
    const dim3 grid((p.n + kWaves * kOutPerWave - 1) / (kWaves * kOutPerWave), (nView + kLanes - 1) / kLanes);
    const size_t lds = static_cast<size_t>(2 * p.n - 1) * sizeof(double);
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
