// Radix-2/4 Stockham FFT in LDS and the two-view FFT filtering body, shared by
// the single-slice filter kernel (filter.hip) and the batched persistent
// kernel (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace ctsim {
namespace dev {

// Aligned to its size, so that LDS accesses are single ds_read_b64 /
// ds_write_b64 (b128 for FP64): at 4-byte alignment each access splits into
// ds_read2_b32 / ds_write2_b32, which a probe kernel timed at 8.3 and 8.1
// cycles per wave instruction against 3.1 and 6.4 for the 8-byte forms.
template <class T>
struct alignas(2 * sizeof(T)) Cplx {
  T x, y;
};

template <class T>
__device__ __forceinline__ Cplx<T> cmul(Cplx<T> a, Cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// The same product with one rounding less per part (the compile-time FP32
// transforms): two multiplies and two fused multiply-adds.
__device__ __forceinline__ Cplx<float> cmul_fma(Cplx<float> a, Cplx<float> b) {
  return {__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x)};
}

// In-place radix-2 Stockham pass over z[0..N) held in LDS.  tw[k] =
// exp(-2 pi i k / N) for k < N/2; dir = -1 forward, +1 inverse (conjugate
// twiddles).  Every lane first reads its butterflies, then the block
// synchronises, then writes: the autosort permutation makes the passes
// write-after-read hazards otherwise.  R = ceil(N / 2 / 256) butterflies per
// lane.
template <class T, int R>
__device__ __forceinline__ void stockham(Cplx<T>* z, const Cplx<T>* tw, int N, int logN, int dir) {
  const int half = N >> 1;
  for (int s = 0, Ls = 1; s < logN; s++, Ls <<= 1) {
    Cplx<T> ya[R], yb[R];
    int idx[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int t = threadIdx.x + r * blockDim.x;
      idx[r] = -1;
      if (t < half) {
        const int k = t & (Ls - 1);
        Cplx<T> a = z[t], b = z[t + half];
        Cplx<T> w = tw[k * (half / Ls)];
        if (dir > 0) w.y = -w.y;
        const Cplx<T> wb = cmul(w, b);
        ya[r] = {a.x + wb.x, a.y + wb.y};
        yb[r] = {a.x - wb.x, a.y - wb.y};
        idx[r] = ((t - k) << 1) + k;
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (idx[r] >= 0) {
        z[idx[r]] = ya[r];
        z[idx[r] + Ls] = yb[r];
      }
    }
    __syncthreads();
  }
}

// Twiddle w^m = exp(dir * 2 pi i m / N) for 0 <= m < N from the half table
// (w^(m + N/2) = -w^m).
template <class T>
__device__ __forceinline__ Cplx<T> twiddle(const Cplx<T>* tw, int m, int half, int dir) {
  Cplx<T> w = tw[m < half ? m : m - half];
  if (m >= half) { w.x = -w.x; w.y = -w.y; }
  if (dir > 0) w.y = -w.y;
  return w;
}

// The same transform as stockham() with radix-4 passes (one radix-2 pass
// first when log2 N is odd): half the LDS round trips and barriers.
// Radix-R Stockham step for butterfly t: k = t mod Ls, inputs z[t + r N/R]
// times w^(r k N/(R Ls)), an R-point DFT, outputs at (t - k) R + k + r Ls.
template <class T, int R>
__device__ __forceinline__ void stockham4(Cplx<T>* z, const Cplx<T>* tw, int N, int logN, int dir) {
  const int half = N >> 1, quarter = N >> 2;
  int Ls = 1, s = 0;
  if (logN & 1) {
    Cplx<T> ya[R], yb[R];
    int idx[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int t = threadIdx.x + r * blockDim.x;
      idx[r] = -1;
      if (t < half) {
        const Cplx<T> a = z[t], b = z[t + half];
        ya[r] = {a.x + b.x, a.y + b.y};
        yb[r] = {a.x - b.x, a.y - b.y};
        idx[r] = t << 1;
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (idx[r] >= 0) {
        z[idx[r]] = ya[r];
        z[idx[r] + 1] = yb[r];
      }
    }
    __syncthreads();
    Ls = 2;
    s = 1;
  }
  constexpr int R4 = (R + 1) / 2;  // radix-4 butterflies per lane
  for (; s < logN; s += 2, Ls <<= 2) {
    Cplx<T> y0[R4], y1[R4], y2[R4], y3[R4];
    int idx[R4];
    const int stride = quarter / Ls;  // N / (4 Ls)
#pragma unroll
    for (int r = 0; r < R4; r++) {
      const int t = threadIdx.x + r * blockDim.x;
      idx[r] = -1;
      if (t < quarter) {
        const int k = t & (Ls - 1);
        Cplx<T> a0 = z[t], a1 = z[t + quarter], a2 = z[t + 2 * quarter], a3 = z[t + 3 * quarter];
        if (Ls > 1) {
          a1 = cmul(twiddle(tw, k * stride, half, dir), a1);
          a2 = cmul(twiddle(tw, 2 * k * stride, half, dir), a2);
          a3 = cmul(twiddle(tw, 3 * k * stride, half, dir), a3);
        }
        // 4-point DFT; the forward transform multiplies a1 - a3 by -i
        const Cplx<T> s02 = {a0.x + a2.x, a0.y + a2.y}, d02 = {a0.x - a2.x, a0.y - a2.y};
        const Cplx<T> s13 = {a1.x + a3.x, a1.y + a3.y}, d13 = {a1.x - a3.x, a1.y - a3.y};
        const Cplx<T> jd = dir < 0 ? Cplx<T>{d13.y, -d13.x} : Cplx<T>{-d13.y, d13.x};
        y0[r] = {s02.x + s13.x, s02.y + s13.y};
        y2[r] = {s02.x - s13.x, s02.y - s13.y};
        y1[r] = {d02.x + jd.x, d02.y + jd.y};
        y3[r] = {d02.x - jd.x, d02.y - jd.y};
        idx[r] = ((t - k) << 2) + k;
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R4; r++) {
      if (idx[r] >= 0) {
        z[idx[r]] = y0[r];
        z[idx[r] + Ls] = y1[r];
        z[idx[r] + 2 * Ls] = y2[r];
        z[idx[r] + 3 * Ls] = y3[r];
      }
    }
    __syncthreads();
  }
}

// Mixed-radix Stockham FFT for FP32 transforms up to 2048 points: one
// radix-2 or radix-4 pass first when log2 N is not a multiple of 3, then
// radix-8 passes (N = 2048: 4, 8, 8, 8), one butterfly per lane and pass, so
// 4 LDS round trips per transform instead of 6.  Twiddles come from a table
// laid out per pass: the pass with span Ls > 1 holds w^(r k N / (R Ls)),
// r = 1..R-1, k < Ls, at offset + (r - 1) Ls + k (offsets in pass order, at
// most N - 4 entries); lane k of a pass then reads consecutive entries (no
// bank conflicts; lanes sharing k broadcast), where the half table's index
// r k N / (R Ls) strides across one bank.
__host__ __device__ constexpr int fft8_first_radix(int logN) {
  return logN % 3 == 1 ? 2 : (logN % 3 == 2 ? 4 : 8);
}

template <class T, int PER_LANE = 8>
__device__ __forceinline__ void pass_twiddles8(const Cplx<T>* __restrict__ twg, Cplx<T>* tw8, int N,
                                               int logN) {
  // the radix-8 passes with twiddles start at Ls = the first radix (a
  // radix-8 first pass has Ls = 1 and needs none, the next has Ls = 8); at
  // most N - 4 entries, PER_LANE per lane (8 up to 2048 points, 16 for
  // 4096), all gathered before the stores
  const int half = N >> 1;
  const int l0 = fft8_first_radix(logN) == 2 ? 1 : (fft8_first_radix(logN) == 4 ? 2 : 3);
  int total = 0;
  for (int lg = l0; lg < logN; lg += 3) total += 7 << lg;
  Cplx<T> v[PER_LANE];
  int dst[PER_LANE];
#pragma unroll
  for (int i = 0; i < PER_LANE; i++) {
    const int e = threadIdx.x + i * blockDim.x;
    dst[i] = -1;
    if (e < total) {
      int lg = l0, off = 0;
      while (e >= off + (7 << lg)) {  // at most three steps
        off += 7 << lg;
        lg += 3;
      }
      const int el = e - off, r = (el >> lg) + 1, k = el & ((1 << lg) - 1);
      v[i] = twiddle(twg, (r * k) << (logN - 3 - lg), half, -1);  // r k N / (8 Ls)
      dst[i] = e;
    }
  }
#pragma unroll
  for (int i = 0; i < PER_LANE; i++)
    if (dst[i] >= 0) tw8[dst[i]] = v[i];
}

// 4-point DFT in place (c1 - c3 times -i forward, +i inverse).
template <class T>
__device__ __forceinline__ void dft4(Cplx<T>& c0, Cplx<T>& c1, Cplx<T>& c2, Cplx<T>& c3, int dir) {
  const Cplx<T> s02 = {c0.x + c2.x, c0.y + c2.y}, d02 = {c0.x - c2.x, c0.y - c2.y};
  const Cplx<T> s13 = {c1.x + c3.x, c1.y + c3.y}, d13 = {c1.x - c3.x, c1.y - c3.y};
  const Cplx<T> jd = dir < 0 ? Cplx<T>{d13.y, -d13.x} : Cplx<T>{-d13.y, d13.x};
  c0 = {s02.x + s13.x, s02.y + s13.y};
  c2 = {s02.x - s13.x, s02.y - s13.y};
  c1 = {d02.x + jd.x, d02.y + jd.y};
  c3 = {d02.x - jd.x, d02.y - jd.y};
}

// R-point DFT of a[0..R) in place, R = 2, 4, 8 (radix 8 by decimation in
// frequency: even outputs are the 4-point DFT of a_r + a_{r+4}, odd outputs
// that of (a_r - a_{r+4}) w8^r).
template <class T, int R>
__device__ __forceinline__ void dft_small(Cplx<T>* a, int dir) {
  if constexpr (R == 2) {
    const Cplx<T> s = {a[0].x + a[1].x, a[0].y + a[1].y}, d = {a[0].x - a[1].x, a[0].y - a[1].y};
    a[0] = s;
    a[1] = d;
  } else if constexpr (R == 4) {
    dft4(a[0], a[1], a[2], a[3], dir);
  } else {
    const T h = T(0.70710678118654752440);
    Cplx<T> b[4], c[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      b[r] = {a[r].x + a[r + 4].x, a[r].y + a[r + 4].y};
      c[r] = {a[r].x - a[r + 4].x, a[r].y - a[r + 4].y};
    }
    // c_r *= w8^r, w8 = (1 + dir i) / sqrt 2
    const T sg = dir < 0 ? T(-1) : T(1);
    c[1] = {h * (c[1].x - sg * c[1].y), h * (c[1].y + sg * c[1].x)};
    c[2] = {-sg * c[2].y, sg * c[2].x};
    c[3] = {-h * (c[3].x + sg * c[3].y), h * (sg * c[3].x - c[3].y)};
    dft4(b[0], b[1], b[2], b[3], dir);
    dft4(c[0], c[1], c[2], c[3], dir);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      a[2 * k] = b[k];
      a[2 * k + 1] = c[k];
    }
  }
}

// One radix-R Stockham pass over z[0..N): butterfly t < N/R (NB per lane)
// reads z[t + r N/R], multiplies input r by tw[off + (r - 1) Ls + k]
// (k = t mod Ls, when Ls > 1), takes the R-point DFT and writes output q to
// (t - k) R + k + q Ls.  Reads, barrier, writes, barrier.
template <class T, int R, int NB>
__device__ __forceinline__ void radix_pass(Cplx<T>* z, const Cplx<T>* tw, int N, int Ls, int off, int dir) {
  const int nb = N / R;
  Cplx<T> y[NB][R];
  int idx[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const int t = threadIdx.x + b * blockDim.x;
    idx[b] = -1;
    if (t < nb) {
      const int k = t & (Ls - 1);
#pragma unroll
      for (int r = 0; r < R; r++) y[b][r] = z[t + r * nb];
      if (Ls > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) {
          Cplx<T> w = tw[off + (r - 1) * Ls + k];
          if (dir > 0) w.y = -w.y;
          y[b][r] = cmul(w, y[b][r]);
        }
      }
      dft_small<T, R>(y[b], dir);
      idx[b] = (t - k) * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; b++) {
    if (idx[b] >= 0) {
#pragma unroll
      for (int q = 0; q < R; q++) z[idx[b] + q * Ls] = y[b][q];
    }
  }
  __syncthreads();
}

// The transform (N <= 2048, 256 lanes) over the table of pass_twiddles8.
template <class T>
__device__ __forceinline__ void stockham8(Cplx<T>* z, const Cplx<T>* tw8, int N, int logN, int dir) {
  int Ls = 1, off = 0;
  const int first = fft8_first_radix(logN);
  if (first == 2) {
    radix_pass<T, 2, 4>(z, tw8, N, 1, 0, dir);
    Ls = 2;
  } else if (first == 4) {
    radix_pass<T, 4, 2>(z, tw8, N, 1, 0, dir);
    Ls = 4;
  }
  for (; Ls < N; Ls <<= 3) {
    radix_pass<T, 8, 1>(z, tw8, N, Ls, off, dir);
    if (Ls > 1) off += 7 * Ls;
  }
}

// Filters rows va and vb (vb < 0: none) of `in` (n cells each, stride n)
// with one complex transform z = x_a + i x_b of length N (a power of two
// >= n): optional FP64 pre-weights w, forward FFT, the two real spectra split
// by Hermitian symmetry, real and imaginary parts multiplied by hre[k] and
// him[k] (k <= N/2, 1/N folded in), recombined, inverse FFT.  On return the
// filtered rows are z[j].x (va) and z[j].y (vb), j < n, in LDS; z and tw
// are LDS arrays of the caller, tw of fft_tw_entries<T>(N) twiddles: FP32
// transforms up to 2048 points take the mixed-radix transform and its
// per-pass table (pass_twiddles8),
// longer ones and FP64 ones the half table (N/2 entries), which keeps a
// 4096-point FP32 or a 2048-point FP64 transform within 48 KB of LDS.
template <class T>
__host__ __device__ constexpr bool fft_per_pass(int N) {
  return sizeof(T) == 4 && N <= 2048;
}
template <class T>
__host__ __device__ constexpr int fft_tw_entries(int N) {
  return fft_per_pass<T>(N) ? N : N / 2;
}

template <class T, int R>
__device__ __forceinline__ void fft_filter_pair(const float* __restrict__ in, const double* __restrict__ w,
                                const T* __restrict__ hre, const T* __restrict__ him,
                                const Cplx<T>* __restrict__ twg, int va, int vb, int n, int N,
                                int logN, Cplx<T>* z, Cplx<T>* tw) {
  // Every global read of the pair is issued before the first LDS store (a
  // loop of load-then-store pays the L2 latency once per iteration): 256
  // lanes, IT = N/256 input cells and HT spectrum multipliers per lane.
  constexpr int IT = 2 * R, HT = R + 1;
  const bool perPass = fft_per_pass<T>(N);
  float xa[IT], xb[IT];
  double wj[IT];
#pragma unroll
  for (int i = 0; i < IT; i++) {
    const int j = threadIdx.x + i * blockDim.x;
    xa[i] = xb[i] = 0.f;
    wj[i] = 1.0;
    if (j < n) {
      xa[i] = in[static_cast<size_t>(va) * n + j];
      if (vb >= 0) xb[i] = in[static_cast<size_t>(vb) * n + j];
      if (w) wj[i] = w[j];
    }
  }
  T mr[HT], mi[HT];
#pragma unroll
  for (int i = 0; i < HT; i++) {
    const int k = threadIdx.x + i * blockDim.x;
    mr[i] = mi[i] = T(0);
    if (k <= N / 2) {
      mr[i] = hre[k];
      mi[i] = him[k];
    }
  }
  if (perPass) pass_twiddles8(twg, tw, N, logN);
  else for (int k = threadIdx.x; k < N / 2; k += blockDim.x) tw[k] = twg[k];
#pragma unroll
  for (int i = 0; i < IT; i++) {
    const int j = threadIdx.x + i * blockDim.x;
    if (j < N) {
      double da = static_cast<double>(xa[i]), db = static_cast<double>(xb[i]);
      if (w) {
        da *= wj[i];
        db *= wj[i];
      }
      z[j] = {static_cast<T>(da), static_cast<T>(db)};
    }
  }
  __syncthreads();
  if (perPass) stockham8<T>(z, tw, N, logN, -1);
  else stockham4<T, R>(z, tw, N, logN, -1);
  // Split the two real spectra, apply the multipliers, recombine.
#pragma unroll
  for (int i = 0; i < HT; i++) {
    const int k = threadIdx.x + i * blockDim.x;
    if (k > N / 2) break;
    const int kn = (N - k) & (N - 1);
    const Cplx<T> zk = z[k], zn = z[kn];
    // X_a = (Z_k + conj Z_{N-k}) / 2, X_b = (Z_k - conj Z_{N-k}) / (2i)
    const T ar = (zk.x + zn.x) * T(0.5), ai = (zk.y - zn.y) * T(0.5);
    const T br = (zk.y + zn.y) * T(0.5), bi = (zn.x - zk.x) * T(0.5);
    const T hr = mr[i], hi = mi[i];
    const T yar = hr * ar, yai = hi * ai, ybr = hr * br, ybi = hi * bi;
    // W_k = Y_a + i Y_b, W_{N-k} = conj(Y_a) + i conj(Y_b)
    const Cplx<T> wk = {yar - ybi, yai + ybr};
    const Cplx<T> wn = {yar + ybi, ybr - yai};
    // pairs (k, N-k) for k <= N/2 are disjoint across lanes: no barrier needed
    z[k] = wk;
    if (kn != k) z[kn] = wn;
  }
  __syncthreads();
  if (perPass) stockham8<T>(z, tw, N, logN, +1);
  else stockham4<T, R>(z, tw, N, logN, +1);
}

// ---------------------------------------------------------------------
// The FP32 transform with the length and the direction known at compile time
// (N = 2^LOGN <= 2048, 256 lanes): the pass sequence unrolls, LDS offsets
// become immediates, and no lane bound or direction select is evaluated.
// Same passes, tables and arithmetic as stockham8.
//
// Element i of these transforms lives in LDS slot fft_slot(i): the low four
// index bits XORed with bits 4-7.  The butterfly outputs of the early passes
// go to indices R apart (R = 2, 4, 8), which in the plain layout fall on a
// few of the banks; the XOR spreads each run of 16 such writes over all of
// them.  Reads at t + r N/R keep their immediate offsets (N/R is a multiple
// of 256 from 512 points up).  Callers read the result through fft_slot.
__host__ __device__ constexpr int fft_slot(int i) { return i ^ ((i >> 4) & 15); }

// Lengths 2^6 ... 2^12 take these transforms; their LDS holds the N points
// and, at 4096 points, the N-entry per-pass twiddle table (64 KB); up to 2048
// points the twiddles are held in registers (RegTwiddles).
__host__ __device__ constexpr bool fft_c_length(int logN) { return logN >= 6 && logN <= 12; }
__host__ __device__ constexpr size_t fft_c_lds(int N) {
  return (N <= 2048 ? 1 : 2) * static_cast<size_t>(N) * sizeof(Cplx<float>);
}

// FP32 twiddle tables for the transforms: the half table exp(-2 pi i m / N),
// m < N / 2 (N floats), followed, for 64 <= N <= 2048, by the radix-8
// passes' twiddles in pass order: w^(r k N / (8 Ls)) at off + (r - 1) Ls + k
// for each pass with span Ls > 1 (r = 1..7, k < Ls), copied from the half
// table with the sign of w^(m + N/2) = -w^m.  A lane's seven twiddles of a
// pass are then consecutive table runs across the wave (the half table's
// index r k strides up to 7 entries per lane).
inline void fft_append_pass_twiddles(std::vector<float>& t32, int N) {
  int logN = 0;
  while ((1 << logN) < N) logN++;
  if (logN < 6 || logN > 11) return;
  const int half = N / 2;
  const int l0 = fft8_first_radix(logN) == 2 ? 1 : (fft8_first_radix(logN) == 4 ? 2 : 3);
  for (int lg = l0; lg < logN; lg += 3)
    for (int r = 1; r < 8; r++)
      for (int k = 0; k < (1 << lg); k++) {
        const int m = (r * k) << (logN - 3 - lg);
        const int e = m < half ? m : m - half;
        const float sg = m < half ? 1.f : -1.f;
        t32.push_back(sg * t32[2 * e]);
        t32.push_back(sg * t32[2 * e + 1]);
      }
}

// Twiddles of the compile-time transforms up to 2048 points in registers:
// each radix-8 pass with span Ls > 1 has at most one butterfly per lane
// (N / 8 <= 256), which needs w^(r k N / (8 Ls)), r = 1..7, k = lane mod Ls;
// the forward transform uses them and the inverse their conjugates.  Loaded
// once from the pass-order table (fft_append_pass_twiddles), they keep the
// twiddle table out of LDS and its reads out of the passes.
template <int LOGN>
struct RegTwiddles {
  static constexpr int kL0 = fft8_first_radix(LOGN) == 2 ? 1 : (fft8_first_radix(LOGN) == 4 ? 2 : 3);
  static constexpr int kPasses = LOGN > kL0 ? (LOGN - kL0 + 2) / 3 : 0;
  Cplx<float> w[kPasses > 0 ? kPasses : 1][7];
  // twg: the half table with the pass-order table after it
  __device__ __forceinline__ void load(const Cplx<float>* __restrict__ twg) {
    const Cplx<float>* pt = twg + (1 << LOGN) / 2;
    int off = 0;
#pragma unroll
    for (int p = 0; p < kPasses; p++) {
      const int lg = kL0 + 3 * p;
      const int k = threadIdx.x & ((1 << lg) - 1);
#pragma unroll
      for (int r = 1; r < 8; r++) w[p][r - 1] = pt[off + ((r - 1) << lg) + k];
      off += 7 << lg;
    }
  }
};

template <int DIR, class T>
__device__ __forceinline__ void dft4c(Cplx<T>& c0, Cplx<T>& c1, Cplx<T>& c2, Cplx<T>& c3) {
  const Cplx<T> s02 = {c0.x + c2.x, c0.y + c2.y}, d02 = {c0.x - c2.x, c0.y - c2.y};
  const Cplx<T> s13 = {c1.x + c3.x, c1.y + c3.y}, d13 = {c1.x - c3.x, c1.y - c3.y};
  const Cplx<T> jd = DIR < 0 ? Cplx<T>{d13.y, -d13.x} : Cplx<T>{-d13.y, d13.x};
  c0 = {s02.x + s13.x, s02.y + s13.y};
  c2 = {s02.x - s13.x, s02.y - s13.y};
  c1 = {d02.x + jd.x, d02.y + jd.y};
  c3 = {d02.x - jd.x, d02.y - jd.y};
}

template <int DIR, int R, class T>
__device__ __forceinline__ void dft_smallc(Cplx<T>* a) {
  if constexpr (R == 2) {
    const Cplx<T> s = {a[0].x + a[1].x, a[0].y + a[1].y}, d = {a[0].x - a[1].x, a[0].y - a[1].y};
    a[0] = s;
    a[1] = d;
  } else if constexpr (R == 4) {
    dft4c<DIR>(a[0], a[1], a[2], a[3]);
  } else {
    const T h = T(0.70710678118654752440);
    constexpr T sg = DIR < 0 ? T(-1) : T(1);
    Cplx<T> b[4], c[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      b[r] = {a[r].x + a[r + 4].x, a[r].y + a[r + 4].y};
      c[r] = {a[r].x - a[r + 4].x, a[r].y - a[r + 4].y};
    }
    c[1] = {h * (c[1].x - sg * c[1].y), h * (c[1].y + sg * c[1].x)};
    c[2] = {-sg * c[2].y, sg * c[2].x};
    c[3] = {-h * (c[3].x + sg * c[3].y), h * (sg * c[3].x - c[3].y)};
    dft4c<DIR>(b[0], b[1], b[2], b[3]);
    dft4c<DIR>(c[0], c[1], c[2], c[3]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      a[2 * k] = b[k];
      a[2 * k + 1] = c[k];
    }
  }
}

// HALF (the first pass only, Ls = 1): inputs N/2 .. N - 1 are zeros and are
// not read (a butterfly's inputs r >= R/2 lie there)
template <int DIR, int R, int LOGN, int LG, int OFF, class T, bool HALF = false>
__device__ __forceinline__ void radix_passc(Cplx<T>* z, const Cplx<T>* tw, const RegTwiddles<LOGN>* rt) {
  static_assert(!HALF || LG == 0, "zero upper half: first pass only");
  constexpr int N = 1 << LOGN, Ls = 1 << LG, nb = N / R;
  constexpr int NB = nb >= 256 ? nb / 256 : 1;
  Cplx<T> y[NB][R];
  int idx[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const int t = threadIdx.x + b * 256;
    idx[b] = -1;
    if (nb >= 256 || t < nb) {
      const int k = t & (Ls - 1);
#pragma unroll
      for (int r = 0; r < R; r++) y[b][r] = (HALF && r >= R / 2) ? Cplx<T>{T(0), T(0)} : z[fft_slot(t + r * nb)];
      if constexpr (Ls > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) {
          Cplx<T> w;
          if constexpr (LOGN <= 11) w = rt->w[(LG - RegTwiddles<LOGN>::kL0) / 3][r - 1];
          else w = tw[OFF + (r - 1) * Ls + k];
          if constexpr (DIR > 0) w.y = -w.y;
          y[b][r] = cmul_fma(w, y[b][r]);
        }
      }
      dft_smallc<DIR, R>(y[b]);
      idx[b] = (t - k) * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; b++) {
    if (nb >= 256 || idx[b] >= 0) {
#pragma unroll
      for (int q = 0; q < R; q++) z[fft_slot(idx[b] + q * Ls)] = y[b][q];
    }
  }
  __syncthreads();
}

// radix-8 passes with span 2^LG, 2^(LG+3), ... < N; table offset OFF
template <int DIR, int LOGN, int LG, int OFF, class T>
__device__ __forceinline__ void radix8_passesc(Cplx<T>* z, const Cplx<T>* tw, const RegTwiddles<LOGN>* rt) {
  if constexpr (LG < LOGN) {
    radix_passc<DIR, 8, LOGN, LG, OFF>(z, tw, rt);
    radix8_passesc<DIR, LOGN, LG + 3, (LG > 0 ? OFF + 7 * (1 << LG) : OFF)>(z, tw, rt);
  }
}

// tw8: the LDS table (4096 points) or, up to 2048 points, unused (rt);
// HALF: the upper half of the input is zero and was not stored
template <int DIR, int LOGN, class T, bool HALF = false>
__device__ __forceinline__ void stockham8c(Cplx<T>* z, const Cplx<T>* tw8, const RegTwiddles<LOGN>* rt) {
  constexpr int first = fft8_first_radix(LOGN);
  if constexpr (first == 2) {
    radix_passc<DIR, 2, LOGN, 0, 0, T, HALF>(z, tw8, rt);
    radix8_passesc<DIR, LOGN, 1, 0>(z, tw8, rt);
  } else if constexpr (first == 4) {
    radix_passc<DIR, 4, LOGN, 0, 0, T, HALF>(z, tw8, rt);
    radix8_passesc<DIR, LOGN, 2, 0>(z, tw8, rt);
  } else {
    radix_passc<DIR, 8, LOGN, 0, 0, T, HALF>(z, tw8, rt);
    radix8_passesc<DIR, LOGN, 3, 0>(z, tw8, rt);
  }
}

// fft_filter_pair for FP32 with N = 2^LOGN <= 4096 at compile time (at 4096
// points tw holds the N-entry table; up to 2048 tw is unused); filtered cell
// j of the rows is z[fft_slot(j)].x (va) and .y (vb).
template <int LOGN>
__device__ __forceinline__ void fft_filter_pair_c(const float* __restrict__ in, const double* __restrict__ w,
                                                  const float* __restrict__ hre, const float* __restrict__ him,
                                                  const Cplx<float>* __restrict__ twg, int va, int vb, int n,
                                                  Cplx<float>* z, Cplx<float>* tw) {
  static_assert(LOGN >= 1 && LOGN <= 12, "FP32 compile-time transforms up to 4096 points");
  constexpr int N = 1 << LOGN;
  constexpr int IT = N >= 256 ? N / 256 : 1, HT = (N / 2 + 1 + 255) / 256;
  float xa[IT], xb[IT];
  double wj[IT];
#pragma unroll
  for (int i = 0; i < IT; i++) {
    const int j = threadIdx.x + i * 256;
    xa[i] = xb[i] = 0.f;
    wj[i] = 1.0;
    if (j < n) {
      xa[i] = in[static_cast<size_t>(va) * n + j];
      if (vb >= 0) xb[i] = in[static_cast<size_t>(vb) * n + j];
      if (w) wj[i] = w[j];
    }
  }
  float mr[HT], mi[HT];
#pragma unroll
  for (int i = 0; i < HT; i++) {
    const int k = threadIdx.x + i * 256;
    mr[i] = mi[i] = 0.f;
    if (k <= N / 2) {
      mr[i] = hre[k];
      mi[i] = him[k];
    }
  }
  RegTwiddles<LOGN> rt;
  if constexpr (LOGN <= 11) rt.load(twg);
  else pass_twiddles8<float, 16>(twg, tw, N, LOGN);
  // rows of at most N/2 cells (the usual zero padding to twice the row):
  // the zero upper half is neither stored nor read by the first pass
  const bool half = n <= N / 2;
#pragma unroll
  for (int i = 0; i < IT; i++) {
    const int j = threadIdx.x + i * 256;
    if ((N >= 256 || j < N) && (j < N / 2 || !half)) {
      double da = static_cast<double>(xa[i]), db = static_cast<double>(xb[i]);
      if (w) {
        da *= wj[i];
        db *= wj[i];
      }
      z[fft_slot(j)] = {static_cast<float>(da), static_cast<float>(db)};
    }
  }
  __syncthreads();
  if (half) stockham8c<-1, LOGN, float, true>(z, tw, &rt);
  else stockham8c<-1, LOGN>(z, tw, &rt);
#pragma unroll
  for (int i = 0; i < HT; i++) {
    const int k = threadIdx.x + i * 256;
    if (k > N / 2) break;
    const int kn = (N - k) & (N - 1);
    const Cplx<float> zk = z[fft_slot(k)], zn = z[fft_slot(kn)];
    const float ar = (zk.x + zn.x) * 0.5f, ai = (zk.y - zn.y) * 0.5f;
    const float br = (zk.y + zn.y) * 0.5f, bi = (zn.x - zk.x) * 0.5f;
    const float hr = mr[i], hi = mi[i];
    const float yar = hr * ar, yai = hi * ai, ybr = hr * br, ybi = hi * bi;
    const Cplx<float> wk = {yar - ybi, yai + ybr};
    const Cplx<float> wn = {yar + ybi, ybr - yai};
    z[fft_slot(k)] = wk;
    if (kn != k) z[fft_slot(kn)] = wn;
  }
  __syncthreads();
  stockham8c<+1, LOGN>(z, tw, &rt);
}

}  // namespace dev
}  // namespace ctsim
