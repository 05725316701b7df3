// Radix-2/4 Stockham FFT in LDS and the two-view FFT filtering body, shared by
// the single-slice filter kernel (filter.hip) and the batched persistent
// kernel (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace ctsim {
namespace dev {

template <class T>
struct Cplx {
  T x, y;
};

template <class T>
__device__ __forceinline__ Cplx<T> cmul(Cplx<T> a, Cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// In-place radix-2 Stockham pass over z[0..N) held in LDS.  tw[k] =
// exp(-2 pi i k / N) for k < N/2; dir = -1 forward, +1 inverse (conjugate
// twiddles).  Every lane first reads its butterflies, then the block
// synchronises, then writes: the autosort permutation makes the passes
// write-after-read hazards otherwise.  R = ceil(N / 2 / 256) butterflies per
// lane.
template <class T, int R>
__device__ void stockham(Cplx<T>* z, const Cplx<T>* tw, int N, int logN, int dir) {
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
__device__ void stockham4(Cplx<T>* z, const Cplx<T>* tw, int N, int logN, int dir) {
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

// Twiddles laid out per radix-4 pass (the passes of stockham4 with Ls > 1):
// the pass with span Ls holds w^(j k N / (4 Ls)), j = 1, 2, 3, k < Ls, at
// offset + (j - 1) Ls + k, offsets in pass order.  Lane k of a pass then
// reads consecutive entries (no bank conflicts; lanes sharing k broadcast),
// where the half table's index j k N / (4 Ls) strides across one bank.  At
// most N - 2 entries.
template <class T>
__device__ void pass_twiddles(const Cplx<T>* __restrict__ twg, Cplx<T>* tw4, int N, int logN) {
  const int half = N >> 1, quarter = N >> 2;
  int off = 0;
  for (int Ls = (logN & 1) ? 2 : 4; Ls <= quarter; Ls <<= 2) {
    const int stride = quarter / Ls;
    for (int e = threadIdx.x; e < 3 * Ls; e += blockDim.x) {
      const int j = e / Ls + 1, k = e - (j - 1) * Ls;
      tw4[off + e] = twiddle(twg, j * k * stride, half, -1);
    }
    off += 3 * Ls;
  }
}

// stockham4 over the per-pass twiddle table of pass_twiddles.
template <class T, int R>
__device__ void stockham4p(Cplx<T>* z, const Cplx<T>* tw4, int N, int logN, int dir) {
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
  constexpr int R4 = (R + 1) / 2;
  int off = 0;
  for (; s < logN; s += 2, Ls <<= 2) {
    Cplx<T> y0[R4], y1[R4], y2[R4], y3[R4];
    int idx[R4];
#pragma unroll
    for (int r = 0; r < R4; r++) {
      const int t = threadIdx.x + r * blockDim.x;
      idx[r] = -1;
      if (t < quarter) {
        const int k = t & (Ls - 1);
        Cplx<T> a0 = z[t], a1 = z[t + quarter], a2 = z[t + 2 * quarter], a3 = z[t + 3 * quarter];
        if (Ls > 1) {
          Cplx<T> w1 = tw4[off + k], w2 = tw4[off + Ls + k], w3 = tw4[off + 2 * Ls + k];
          if (dir > 0) { w1.y = -w1.y; w2.y = -w2.y; w3.y = -w3.y; }
          a1 = cmul(w1, a1);
          a2 = cmul(w2, a2);
          a3 = cmul(w3, a3);
        }
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
    if (Ls > 1) off += 3 * Ls;
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

// Filters rows va and vb (vb < 0: none) of `in` (n cells each, stride n)
// with one complex transform z = x_a + i x_b of length N (a power of two
// >= n): optional FP64 pre-weights w, forward FFT, the two real spectra split
// by Hermitian symmetry, real and imaginary parts multiplied by hre[k] and
// him[k] (k <= N/2, 1/N folded in), recombined, inverse FFT.  On return the
// filtered rows are z[j].x (va) and z[j].y (vb), j < n, in LDS; z and tw
// are LDS arrays of the caller, tw of fft_tw_entries<T>(N) twiddles: FP32
// transforms up to 2048 points take the per-pass table (pass_twiddles),
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
__device__ void fft_filter_pair(const float* __restrict__ in, const double* __restrict__ w,
                                const T* __restrict__ hre, const T* __restrict__ him,
                                const Cplx<T>* __restrict__ twg, int va, int vb, int n, int N,
                                int logN, Cplx<T>* z, Cplx<T>* tw) {
  const bool perPass = fft_per_pass<T>(N);
  if (perPass) pass_twiddles(twg, tw, N, logN);
  else for (int k = threadIdx.x; k < N / 2; k += blockDim.x) tw[k] = twg[k];
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    T a = 0, b = 0;
    if (j < n) {
      double xa = static_cast<double>(in[static_cast<size_t>(va) * n + j]);
      double xb = vb >= 0 ? static_cast<double>(in[static_cast<size_t>(vb) * n + j]) : 0.0;
      if (w) {
        xa *= w[j];
        xb *= w[j];
      }
      a = static_cast<T>(xa);
      b = static_cast<T>(xb);
    }
    z[j] = {a, b};
  }
  __syncthreads();
  if (perPass) stockham4p<T, R>(z, tw, N, logN, -1);
  else stockham4<T, R>(z, tw, N, logN, -1);
  // Split the two real spectra, apply the multipliers, recombine.
  for (int k = threadIdx.x; k <= N / 2; k += blockDim.x) {
    const int kn = (N - k) & (N - 1);
    const Cplx<T> zk = z[k], zn = z[kn];
    // X_a = (Z_k + conj Z_{N-k}) / 2, X_b = (Z_k - conj Z_{N-k}) / (2i)
    const T ar = (zk.x + zn.x) * T(0.5), ai = (zk.y - zn.y) * T(0.5);
    const T br = (zk.y + zn.y) * T(0.5), bi = (zn.x - zk.x) * T(0.5);
    const T hr = hre[k], hi = him[k];
    const T yar = hr * ar, yai = hi * ai, ybr = hr * br, ybi = hi * bi;
    // W_k = Y_a + i Y_b, W_{N-k} = conj(Y_a) + i conj(Y_b)
    const Cplx<T> wk = {yar - ybi, yai + ybr};
    const Cplx<T> wn = {yar + ybi, ybr - yai};
    // pairs (k, N-k) for k <= N/2 are disjoint across lanes: no barrier needed
    z[k] = wk;
    if (kn != k) z[kn] = wn;
  }
  __syncthreads();
  if (perPass) stockham4p<T, R>(z, tw, N, logN, +1);
  else stockham4<T, R>(z, tw, N, logN, +1);
}

}  // namespace dev
}  // namespace ctsim
