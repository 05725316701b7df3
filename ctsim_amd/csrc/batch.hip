// Batched mixed-precision FBP of many slices of uneven size, detector count,
// view count and geometry in one persistent kernel, replayed from a HIP graph.
//
// The reference reconstructs one slice at a time (tools/pjrec.cpp:342-347,
// Reconstructor::reconstructView, libctsim/reconstruct.cpp:156-239).  Here a
// batch plan holds, per slice, the reference's filter (ProcessSignal::init)
// and backprojector geometry (Backproject ctor), and a work list:
//   FILTER items  two views of one slice: FFT filtering in LDS (convolution
//                 filters as an FFT convolution with the kernel's transform),
//                 then the interval coefficients of the slice's
//                 interpolation, written for its backprojection;
//   BP items      one tile of one slice over all its views (the mixed
//                 backprojection body), image = rotScale * sum over views;
//                 with the rotationally symmetric body, one quadrant tile
//                 and one group of base views, the tile's last group adding
//                 the groups' partial images.
// Slices are dealt to the 8 XCDs by longest-processing-time first, so a
// slice's coefficients are produced and consumed under one L2; a slice that
// costs more than 1/32 of the batch is dealt item by item over all eight.
// Each XCD has its own queue (its filter items, then its BP items, largest
// slices first); workgroups pull items from their XCD's queue with an
// atomic counter.  A BP item waits (bounded spin, error flag on expiry)
// until its slice's filter items are done.  Filter items never wait, and in
// every queue they precede all BP items, so by the time a workgroup takes a
// BP item every filter item of its queue has been taken by a running
// workgroup; each filter item a BP item can wait on (in any queue) is
// therefore finished or held by a running workgroup, and the wait ends.
// Fused mode runs both item kinds in one launch; two-phase mode launches
// the filter items and then the BP items (no waiting, smaller LDS for the
// BP launch).  The autotuner picks the mode, the tile shape, the views per
// LDS chunk and the symmetric body's view groups.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstring>
#include <vector>

#include "bp_plan.h"
#include "fft_device.h"
#include "filter_plan.h"
#include "mixed_device.h"

namespace ctsim {
namespace {

using dev::CoefT;
using dev::Cplx;
using dev::MixedArgs;

constexpr int kFilterItem = 0, kBpItem = 1;
constexpr int kSpinLimit = 1 << 22;  // about 0.2 s of polling

struct SliceDev {
  // filtering
  int n, N, logN, R, nView, nFilterItems;
  const float* sino;
  const double* w;
  const float* hre;
  const float* him;
  const Cplx<float>* tw;
  // backprojection
  int gk, interp, lastIncl, deg;
  int sym, symS, symK;  // rotationally symmetric body (tiles of the quadrant)
  int P;                // pixels per lane of this slice's tiles
  int mirNM;            // plain slices: 2 or 4 runs the mirror body (two-phase mode), 1 the plain one
  MixedArgs bp;  // bp.coef, bp.image, bp.vc set; overwrite = 1, split = 1
};

struct Item {
  int kind, slice, a, pad;
};

// Zero the queue heads, the per-slice counters and the error word (a
// kernel, not a memset, so the graph holds no memset node): every run
// starts clean and reports only its own timeouts.
__global__ void batch_reset_kernel(int* counters, int n, int* error) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) counters[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *error = 0;
}

// The view groups of the symmetric slices added after the backprojection
// (a kernel boundary orders them, where a device-scope fence per workgroup
// would write back the XCD's L2 each time): slice list[blockIdx.y],
// image = rotScale * sum of its group partial images in group order, FP64.
__global__ __launch_bounds__(256) void batch_combine_kernel(const SliceDev* __restrict__ slices,
                                                            const int* __restrict__ list) {
  const dev::MixedArgs& a = slices[list[blockIdx.y]].bp;
  const size_t npix = static_cast<size_t>(a.nx) * a.ny;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < npix;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    double v = 0;
    for (int g = 0; g < a.split; g++) v += a.partial[g * npix + i];
    a.image[i] = static_cast<float>(static_cast<double>(a.endScale) * v);
  }
}

template <int P>
__device__ __forceinline__ void bp_item(const MixedArgs& a, int gk, int interp, int deg, int tile, void* smem,
                        float4* s_v, int* s_lo) {
  using dev::bp_mixed_tile;
  if (interp == INTERP_CUBIC) {
    float4* w = static_cast<float4*>(smem);
    if (gk == 0) bp_mixed_tile<0, INTERP_CUBIC, P, 0>(a, tile, 0, w, s_v, s_lo);
    else if (gk == 2) bp_mixed_tile<2, INTERP_CUBIC, P, 0>(a, tile, 0, w, s_v, s_lo);
    else if (deg == 4) bp_mixed_tile<1, INTERP_CUBIC, P, 4>(a, tile, 0, w, s_v, s_lo);
    else if (deg == 6) bp_mixed_tile<1, INTERP_CUBIC, P, 6>(a, tile, 0, w, s_v, s_lo);
    else if (deg == 8) bp_mixed_tile<1, INTERP_CUBIC, P, 8>(a, tile, 0, w, s_v, s_lo);
    else bp_mixed_tile<1, INTERP_CUBIC, P, 0>(a, tile, 0, w, s_v, s_lo);
  } else {
    float2* w = static_cast<float2*>(smem);
    if (gk == 0) bp_mixed_tile<0, INTERP_LINEAR, P, 0>(a, tile, 0, w, s_v, s_lo);
    else if (gk == 2) bp_mixed_tile<2, INTERP_LINEAR, P, 0>(a, tile, 0, w, s_v, s_lo);
    else if (deg == 4) bp_mixed_tile<1, INTERP_LINEAR, P, 4>(a, tile, 0, w, s_v, s_lo);
    else if (deg == 6) bp_mixed_tile<1, INTERP_LINEAR, P, 6>(a, tile, 0, w, s_v, s_lo);
    else if (deg == 8) bp_mixed_tile<1, INTERP_LINEAR, P, 8>(a, tile, 0, w, s_v, s_lo);
    else bp_mixed_tile<1, INTERP_LINEAR, P, 0>(a, tile, 0, w, s_v, s_lo);
  }
}

// The mirror body of a plain slice whose views have mirror partners.
template <int P>
__device__ __forceinline__ void bp_mirror_item(const MixedArgs& a, int gk, int interp, int deg, int NM, int tile) {
  using dev::bp_mirror_tile;
  if (interp == INTERP_CUBIC) {
    if (NM == 4) {
      if (gk == 0) bp_mirror_tile<0, INTERP_CUBIC, P, 0, 4>(a, tile, 0);
      else if (gk == 2) bp_mirror_tile<2, INTERP_CUBIC, P, 0, 4>(a, tile, 0);
      else if (deg == 4) bp_mirror_tile<1, INTERP_CUBIC, P, 4, 4>(a, tile, 0);
      else if (deg == 6) bp_mirror_tile<1, INTERP_CUBIC, P, 6, 4>(a, tile, 0);
      else if (deg == 8) bp_mirror_tile<1, INTERP_CUBIC, P, 8, 4>(a, tile, 0);
      else bp_mirror_tile<1, INTERP_CUBIC, P, 0, 4>(a, tile, 0);
    } else {
      if (gk == 0) bp_mirror_tile<0, INTERP_CUBIC, P, 0, 2>(a, tile, 0);
      else if (gk == 2) bp_mirror_tile<2, INTERP_CUBIC, P, 0, 2>(a, tile, 0);
      else if (deg == 4) bp_mirror_tile<1, INTERP_CUBIC, P, 4, 2>(a, tile, 0);
      else if (deg == 6) bp_mirror_tile<1, INTERP_CUBIC, P, 6, 2>(a, tile, 0);
      else if (deg == 8) bp_mirror_tile<1, INTERP_CUBIC, P, 8, 2>(a, tile, 0);
      else bp_mirror_tile<1, INTERP_CUBIC, P, 0, 2>(a, tile, 0);
    }
  } else {
    if (NM == 4) {
      if (gk == 0) bp_mirror_tile<0, INTERP_LINEAR, P, 0, 4>(a, tile, 0);
      else if (gk == 2) bp_mirror_tile<2, INTERP_LINEAR, P, 0, 4>(a, tile, 0);
      else if (deg == 4) bp_mirror_tile<1, INTERP_LINEAR, P, 4, 4>(a, tile, 0);
      else if (deg == 6) bp_mirror_tile<1, INTERP_LINEAR, P, 6, 4>(a, tile, 0);
      else if (deg == 8) bp_mirror_tile<1, INTERP_LINEAR, P, 8, 4>(a, tile, 0);
      else bp_mirror_tile<1, INTERP_LINEAR, P, 0, 4>(a, tile, 0);
    } else {
      if (gk == 0) bp_mirror_tile<0, INTERP_LINEAR, P, 0, 2>(a, tile, 0);
      else if (gk == 2) bp_mirror_tile<2, INTERP_LINEAR, P, 0, 2>(a, tile, 0);
      else if (deg == 4) bp_mirror_tile<1, INTERP_LINEAR, P, 4, 2>(a, tile, 0);
      else if (deg == 6) bp_mirror_tile<1, INTERP_LINEAR, P, 6, 2>(a, tile, 0);
      else if (deg == 8) bp_mirror_tile<1, INTERP_LINEAR, P, 8, 2>(a, tile, 0);
      else bp_mirror_tile<1, INTERP_LINEAR, P, 0, 2>(a, tile, 0);
    }
  }
}

// The symmetric body (software-pipelined) for slices whose views come in
// rotated copies.
template <int GK, int INTERP, int P, int DEG>
__device__ __forceinline__ void bp_sym_sk(const SliceDev& s, int tile, int grp, void* smem, float2* s_ab,
                                          dev::SymWin* s_win) {
  auto* w = static_cast<CoefT<INTERP>*>(smem);
  if (s.symS == 4) dev::bp_sym_tile_pipe<GK, INTERP, P, DEG, 4, 1>(s.bp, tile, grp, w, s_ab, s_win);
  else if (s.symK == 1) dev::bp_sym_tile_pipe<GK, INTERP, P, DEG, 2, 1>(s.bp, tile, grp, w, s_ab, s_win);
  else dev::bp_sym_tile_pipe<GK, INTERP, P, DEG, 2, 2>(s.bp, tile, grp, w, s_ab, s_win);
}

template <int P, int INTERP>
__device__ __forceinline__ void bp_sym_interp(const SliceDev& s, int tile, int grp, void* smem,
                                              float2* s_ab, dev::SymWin* s_win) {
  if (s.gk == 0) bp_sym_sk<0, INTERP, P, 0>(s, tile, grp, smem, s_ab, s_win);
  else if (s.gk == 2) bp_sym_sk<2, INTERP, P, 0>(s, tile, grp, smem, s_ab, s_win);
  else if (s.deg == 4) bp_sym_sk<1, INTERP, P, 4>(s, tile, grp, smem, s_ab, s_win);
  else if (s.deg == 6) bp_sym_sk<1, INTERP, P, 6>(s, tile, grp, smem, s_ab, s_win);
  else if (s.deg == 8) bp_sym_sk<1, INTERP, P, 8>(s, tile, grp, smem, s_ab, s_win);
  else bp_sym_sk<1, INTERP, P, 0>(s, tile, grp, smem, s_ab, s_win);
}

// R > 0: the runtime-length transform with R butterflies per lane; R < 0:
// the compile-time transform of 2^-R points.
template <int R>
__device__ __forceinline__ void filter_item(const SliceDev& s, int pair, Cplx<float>* z, Cplx<float>* tw) {
  const int va = 2 * pair, vb = va + 1 < s.nView ? va + 1 : -1;
  if constexpr (R < 0)
    dev::fft_filter_pair_c<-R>(s.sino, s.w, s.hre, s.him, s.tw, va, vb, s.n, z, tw);
  else
    dev::fft_filter_pair<float, R>(s.sino, s.w, s.hre, s.him, s.tw, va, vb, s.n, s.N, s.logN, z, tw);
  // interval coefficients of both rows straight from LDS
  const int n = s.n;
  // the compile-time transforms leave cell j in slot fft_slot(j)
  auto ra = [z](int j) { return z[R < 0 ? dev::fft_slot(j) : j].x; };
  auto rb = [z](int j) { return z[R < 0 ? dev::fft_slot(j) : j].y; };
  if (s.interp == INTERP_CUBIC) {
    float4* c = static_cast<float4*>(const_cast<void*>(s.bp.coef));
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      c[static_cast<size_t>(va) * n + j] = dev::interval_coef<INTERP_CUBIC, false>(ra, j, n);
      if (vb >= 0) c[static_cast<size_t>(vb) * n + j] = dev::interval_coef<INTERP_CUBIC, false>(rb, j, n);
    }
  } else {
    float2* c = static_cast<float2*>(const_cast<void*>(s.bp.coef));
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      if (s.lastIncl) {
        c[static_cast<size_t>(va) * n + j] = dev::interval_coef<INTERP_LINEAR, true>(ra, j, n);
        if (vb >= 0) c[static_cast<size_t>(vb) * n + j] = dev::interval_coef<INTERP_LINEAR, true>(rb, j, n);
      } else {
        c[static_cast<size_t>(va) * n + j] = dev::interval_coef<INTERP_LINEAR, false>(ra, j, n);
        if (vb >= 0) c[static_cast<size_t>(vb) * n + j] = dev::interval_coef<INTERP_LINEAR, false>(rb, j, n);
      }
    }
  }
}

// counters: [0, 8) queue heads, [8, 8 + nSlices) filter items done per slice.
// SYM instantiates the symmetric body (P <= 2); the plain kernel leaves it
// out and keeps its own register budget.  DO_F / DO_B instantiate the filter
// and the backprojection items: the fused launch takes both, the two-phase
// filter launches only the first (the two-phase backprojection runs in the
// class kernels below, one body each).
template <int P, bool SYM, bool DO_F, bool DO_B, int FL = 0, bool MIR = false>
__device__ __forceinline__ void fbp_batch_body(const SliceDev* __restrict__ slices,
                                               const Item* __restrict__ items,
                                               const int* __restrict__ queueStart, int* counters,
                                               int* error) {
  extern __shared__ float4 smem[];
  // s_meta holds the plain body's per-view constants (kMaxChunk + 2 float4)
  // or the symmetric body's 256 windows and 64 view constants
  static_assert(sizeof(dev::SymWin) == 16, "window record is one float4");
  constexpr int kMeta = SYM ? 288 : dev::kMaxChunk + 2;
  __shared__ float4 s_meta[kMeta];
  __shared__ int s_lo[dev::kMaxChunk];
  __shared__ int s_item;
  float4* s_v = s_meta;
  dev::SymWin* s_win = reinterpret_cast<dev::SymWin*>(s_meta);
  float2* s_ab = reinterpret_cast<float2*>(s_meta + 256);
  const int q = blockIdx.x & 7;
  const int qBegin = queueStart[q], qEnd = queueStart[q + 1];
  int* done = counters + 8;
  while (true) {
    if (threadIdx.x == 0) s_item = qBegin + atomicAdd(&counters[q], 1);
    __syncthreads();
    // wave-uniform, so the item and its slice descriptor load into scalar
    // registers instead of occupying vector registers in every lane
    const int i = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (i >= qEnd) break;  // every workgroup ends when its queue is drained
    const Item it = items[i];
    const SliceDev& s = slices[it.slice];
    if (it.kind == kFilterItem) {
      if constexpr (DO_F) {
        Cplx<float>* z = reinterpret_cast<Cplx<float>*>(smem);
        Cplx<float>* tw = z + s.N;
        if constexpr (FL > 0) {
          filter_item<-FL>(s, it.a, z, tw);  // every slice of this launch has 2^FL points
        } else switch (s.logN) {
          case 6: filter_item<-6>(s, it.a, z, tw); break;
          case 7: filter_item<-7>(s, it.a, z, tw); break;
          case 8: filter_item<-8>(s, it.a, z, tw); break;
          case 9: filter_item<-9>(s, it.a, z, tw); break;
          case 10: filter_item<-10>(s, it.a, z, tw); break;
          case 11: filter_item<-11>(s, it.a, z, tw); break;
          case 12: filter_item<-12>(s, it.a, z, tw); break;
          default:
            switch (s.R) {
              case 1: filter_item<1>(s, it.a, z, tw); break;
              case 2: filter_item<2>(s, it.a, z, tw); break;
              case 4: filter_item<4>(s, it.a, z, tw); break;
              default: filter_item<8>(s, it.a, z, tw); break;
            }
        }
        if constexpr (DO_B) {
          // fused launch: release the coefficients to the backprojection
          // items (a device-scope fence writes back the XCD's L2; the
          // two-phase launches need none, their kernel boundary orders them)
          __threadfence();  // every lane releases its coefficient stores device-wide
          __syncthreads();
          if (threadIdx.x == 0) atomicAdd(&done[it.slice], 1);
        }
      }
    } else if constexpr (DO_B) {
      if constexpr (DO_F) {  // fused launch: wait for the slice's filter items
        if (threadIdx.x == 0) {
          int spins = 0;
          while (__hip_atomic_load(&done[it.slice], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                 s.nFilterItems) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > kSpinLimit) {
              atomicOr(error, 1);
              break;
            }
          }
          __threadfence();  // acquire: the coefficients are visible below
        }
        __syncthreads();
      }
      if constexpr (SYM && P <= 2) {
        if (s.sym) {
          if (s.interp == INTERP_CUBIC) bp_sym_interp<P, INTERP_CUBIC>(s, it.a, it.pad, smem, s_ab, s_win);
          else bp_sym_interp<P, INTERP_LINEAR>(s, it.a, it.pad, smem, s_ab, s_win);
        } else {
          bp_item<P>(s.bp, s.gk, s.interp, s.deg, it.a, smem, s_v, s_lo);
        }
      } else if constexpr (MIR) {
        bp_mirror_item<P>(s.bp, s.gk, s.interp, s.deg, s.mirNM, it.a);
      } else {
        bp_item<P>(s.bp, s.gk, s.interp, s.deg, it.a, smem, s_v, s_lo);
      }
    }
    __syncthreads();  // LDS is reused by the next item
  }
}

// The fused launch and the generic filter launch.
template <int P, bool SYM, bool DO_F, bool DO_B>
__global__ __launch_bounds__(256) void fbp_batch_kernel(const SliceDev* __restrict__ slices,
                                                        const Item* __restrict__ items,
                                                        const int* __restrict__ queueStart,
                                                        int* counters, int* error) {
  fbp_batch_body<P, SYM, DO_F, DO_B>(slices, items, queueStart, counters, error);
}

// The filter launch of the slices whose transforms have 2^FL points (FL in
// 6..11), at that length's register budget.
template <int FL>
__global__ __launch_bounds__(256) void fbp_batch_filter_kernel(const SliceDev* __restrict__ slices,
                                                               const Item* __restrict__ items,
                                                               const int* __restrict__ queueStart,
                                                               int* counters, int* error) {
  fbp_batch_body<1, false, true, false, FL>(slices, items, queueStart, counters, error);
}


// The two-phase backprojection launch of the plain-body items of a batch
// whose slices span several classes: one launch for all of them, dealt over
// the XCD queues as a whole (per-class launches each wait for their most
// loaded queue, which on a mixed batch cost more than the register budget
// of the largest body).
template <int P>
__global__ __launch_bounds__(256) void fbp_batch_bp_kernel(const SliceDev* __restrict__ slices,
                                                           const Item* __restrict__ items,
                                                           const int* __restrict__ queueStart,
                                                           int* counters, int* error) {
  fbp_batch_body<P, false, false, true>(slices, items, queueStart, counters, error);
}

// The two-phase backprojection launch of the mirror-body items of a batch
// whose mirror slices span several classes (one launch, as above).
template <int P>
__global__ __launch_bounds__(256) void fbp_batch_mirror_kernel(const SliceDev* __restrict__ slices,
                                                               const Item* __restrict__ items,
                                                               const int* __restrict__ queueStart,
                                                               int* counters, int* error) {
  fbp_batch_body<P, false, false, true, 0, true>(slices, items, queueStart, counters, error);
}

// The two-phase backprojection launch of the symmetric slices of one class
// (geometry, interpolation, fan polynomial degree, symmetry): one body at its
// own register budget, the single-slice kernel's, where the kernel that
// serves every class holds the largest body's.
template <int P, int GK, int INTERP, int DEG, int S, int K>
__global__ __launch_bounds__(256) void fbp_batch_sym_class_kernel(const SliceDev* __restrict__ slices,
                                                                  const Item* __restrict__ items,
                                                                  const int* __restrict__ queueStart,
                                                                  int* counters, int* /*error: no waits here*/) {
  extern __shared__ float4 smem[];
  __shared__ dev::SymWin s_win[256];
  __shared__ float2 s_ab[64];
  __shared__ int s_item;
  const int q = blockIdx.x & 7;
  const int qBegin = queueStart[q], qEnd = queueStart[q + 1];
  while (true) {
    if (threadIdx.x == 0) s_item = qBegin + atomicAdd(&counters[q], 1);
    __syncthreads();
    const int i = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (i >= qEnd) break;  // every workgroup ends when its queue is drained
    const Item it = items[i];
    const SliceDev& s = slices[it.slice];
    dev::bp_sym_tile_pipe<GK, INTERP, P, DEG, S, K>(s.bp, it.a, it.pad,
                                                    reinterpret_cast<CoefT<INTERP>*>(smem), s_ab, s_win);
    __syncthreads();  // LDS is reused by the next item
  }
}

template <int P, int GK, int INTERP, int DEG>
const void* sym_class_kernel_sk(int S, int K) {
  if (S == 4) return reinterpret_cast<const void*>(fbp_batch_sym_class_kernel<P, GK, INTERP, DEG, 4, 1>);
  if (K == 1) return reinterpret_cast<const void*>(fbp_batch_sym_class_kernel<P, GK, INTERP, DEG, 2, 1>);
  return reinterpret_cast<const void*>(fbp_batch_sym_class_kernel<P, GK, INTERP, DEG, 2, 2>);
}

template <int P, int INTERP>
const void* sym_class_kernel_g(int gk, int deg, int S, int K) {
  if (gk == 0) return sym_class_kernel_sk<P, 0, INTERP, 0>(S, K);
  if (gk == 2) return sym_class_kernel_sk<P, 2, INTERP, 0>(S, K);
  switch (deg) {
    case 4: return sym_class_kernel_sk<P, 1, INTERP, 4>(S, K);
    case 6: return sym_class_kernel_sk<P, 1, INTERP, 6>(S, K);
    case 8: return sym_class_kernel_sk<P, 1, INTERP, 8>(S, K);
    default: return sym_class_kernel_sk<P, 1, INTERP, 0>(S, K);
  }
}

// The two-phase backprojection launch of the plain-body items of one class
// (geometry, interpolation, fan polynomial degree) at that body's register
// budget.
template <int P, int GK, int INTERP, int DEG, int NM>
__global__ __launch_bounds__(256) void fbp_batch_plain_class_kernel(const SliceDev* __restrict__ slices,
                                                                    const Item* __restrict__ items,
                                                                    const int* __restrict__ queueStart,
                                                                    int* counters, int* /*error: no waits here*/) {
  extern __shared__ float4 smem[];
  __shared__ float4 s_v[dev::kMaxChunk + 2];
  __shared__ int s_lo[dev::kMaxChunk];
  __shared__ int s_item;
  const int q = blockIdx.x & 7;
  const int qBegin = queueStart[q], qEnd = queueStart[q + 1];
  while (true) {
    if (threadIdx.x == 0) s_item = qBegin + atomicAdd(&counters[q], 1);
    __syncthreads();
    const int i = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (i >= qEnd) break;  // every workgroup ends when its queue is drained
    const Item it = items[i];
    const SliceDev& s = slices[it.slice];
    if constexpr (NM > 1)
      dev::bp_mirror_tile<GK, INTERP, P, DEG, NM>(s.bp, it.a, 0);
    else
      dev::bp_mixed_tile<GK, INTERP, P, DEG>(s.bp, it.a, 0, reinterpret_cast<CoefT<INTERP>*>(smem), s_v, s_lo);
    __syncthreads();  // LDS is reused by the next item
  }
}

template <int P, int INTERP, int NM>
const void* plain_class_kernel_g(int gk, int deg) {
  if (gk == 0) return reinterpret_cast<const void*>(fbp_batch_plain_class_kernel<P, 0, INTERP, 0, NM>);
  if (gk == 2) return reinterpret_cast<const void*>(fbp_batch_plain_class_kernel<P, 2, INTERP, 0, NM>);
  switch (deg) {
    case 4: return reinterpret_cast<const void*>(fbp_batch_plain_class_kernel<P, 1, INTERP, 4, NM>);
    case 6: return reinterpret_cast<const void*>(fbp_batch_plain_class_kernel<P, 1, INTERP, 6, NM>);
    case 8: return reinterpret_cast<const void*>(fbp_batch_plain_class_kernel<P, 1, INTERP, 8, NM>);
    default: return reinterpret_cast<const void*>(fbp_batch_plain_class_kernel<P, 1, INTERP, 0, NM>);
  }
}

template <int P, int INTERP>
const void* plain_class_kernel_nm(int gk, int deg, int NM) {
  if (NM == 4) return plain_class_kernel_g<P, INTERP, 4>(gk, deg);
  if (NM == 2) return plain_class_kernel_g<P, INTERP, 2>(gk, deg);
  return plain_class_kernel_g<P, INTERP, 1>(gk, deg);
}

// The class kernel of plain-body slices (P 1, 2 or 4; NM 1, or 2 and 4 for
// the mirror body).
const void* plain_class_kernel(int P, int gk, int interp, int deg, int NM) {
  const bool cubic = interp == INTERP_CUBIC;
  if (P == 1)
    return cubic ? plain_class_kernel_nm<1, INTERP_CUBIC>(gk, deg, NM) : plain_class_kernel_nm<1, INTERP_LINEAR>(gk, deg, NM);
  if (P == 2)
    return cubic ? plain_class_kernel_nm<2, INTERP_CUBIC>(gk, deg, NM) : plain_class_kernel_nm<2, INTERP_LINEAR>(gk, deg, NM);
  return cubic ? plain_class_kernel_nm<4, INTERP_CUBIC>(gk, deg, NM) : plain_class_kernel_nm<4, INTERP_LINEAR>(gk, deg, NM);
}

// The class kernel of symmetric slices (P 1 or 2).
const void* sym_class_kernel(int P, int gk, int interp, int deg, int S, int K) {
  if (P == 1)
    return interp == INTERP_CUBIC ? sym_class_kernel_g<1, INTERP_CUBIC>(gk, deg, S, K)
                                  : sym_class_kernel_g<1, INTERP_LINEAR>(gk, deg, S, K);
  return interp == INTERP_CUBIC ? sym_class_kernel_g<2, INTERP_CUBIC>(gk, deg, S, K)
                                : sym_class_kernel_g<2, INTERP_LINEAR>(gk, deg, S, K);
}

// ------------------------------------------------------------------ host

void fft_host(std::vector<std::complex<double>>& a) {
  const int n = static_cast<int>(a.size());
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (int len = 2; len <= n; len <<= 1) {
    for (int i = 0; i < n; i += len)
      for (int k = 0; k < len / 2; k++) {
        const double ang = -2 * kPi * k / len;
        const std::complex<double> w(cos(ang), sin(ang));
        const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
  }
}

}  // namespace
}  // namespace ctsim

using namespace ctsim;

struct ctsim_batch {
  struct Slice {
    FilterPlan fp;
    BpPlan bp;
    std::vector<double> angles;
    int N = 0, logN = 0;
    DevBuf mul, tw;  // FFT-convolution multipliers and twiddles (convolution filters)
    DevBuf vc, coef;
    MirrorViews mir;  // mirror partners of the views (in vc)
    double cost = 0;
  };
  std::vector<Slice> s;
  int P = 2, lanesY = 16, chunkBudget = 32 * 1024;
  bool fused = true;
  // 0: plain body; 1, 2, 4: symmetric body with that many view groups per
  // tile; + 8: the mirror body for the other slices (two-phase mode)
  int sym = 1;
  int ldsFilter = 0;
  std::vector<SliceDev> host;
  DevBuf d_slices, d_itemsF, d_itemsAll, d_qF, d_qAll, d_counters, d_error;
  // two-phase filter launches, one per transform length 2^6 ... 2^11
  // (specialised kernels) and one for the other lengths (logN 0)
  struct FilterClass {
    int logN = 0, qOff = 0, grid = 0;
    size_t lds = 0;
  };
  std::vector<FilterClass> fclass;
  DevBuf d_itemsFc, d_qFc;
  // two-phase backprojection: one launch per class of symmetric slices (body,
  // geometry, interpolation, fan polynomial degree, symmetry, P), and the
  // plain-body slices in one launch of their class kernel when they share a
  // class, else in one generic launch
  struct BpClass {
    const void* kernel = nullptr;
    int qOff = 0, grid = 0;
    size_t lds = 0;
    bool generic = false;  // fbp_batch_bp_kernel (takes the error flag too)
  };
  std::vector<BpClass> bclass;
  DevBuf d_itemsBc, d_qBc;
  DevBuf d_partial;      // view-group partial images of the symmetric slices
  DevBuf d_combine;      // slices whose view groups are added after the backprojection
  int nCombine = 0;
  int gridF = 0, gridAll = 0;
  size_t ldsAll = 0;
  std::vector<const float*> sinos;
  std::vector<float*> images;
  hipStream_t stream = nullptr;
  hipEvent_t evIn = nullptr, evOut = nullptr;
  hipGraphExec_t graph = nullptr;
  bool shapeDirty = true;
  double pvu = 0;
  ~ctsim_batch() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (evIn) (void)hipEventDestroy(evIn);
    if (evOut) (void)hipEventDestroy(evOut);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace ctsim {
namespace {

// The filter-only kernel of one transform length class.
const void* filter_kernel_of(int logN) {
  switch (logN) {
    case 6: return reinterpret_cast<const void*>(fbp_batch_filter_kernel<6>);
    case 7: return reinterpret_cast<const void*>(fbp_batch_filter_kernel<7>);
    case 8: return reinterpret_cast<const void*>(fbp_batch_filter_kernel<8>);
    case 9: return reinterpret_cast<const void*>(fbp_batch_filter_kernel<9>);
    case 10: return reinterpret_cast<const void*>(fbp_batch_filter_kernel<10>);
    case 11: return reinterpret_cast<const void*>(fbp_batch_filter_kernel<11>);
    case 12: return reinterpret_cast<const void*>(fbp_batch_filter_kernel<12>);
    default: return reinterpret_cast<const void*>(fbp_batch_kernel<1, false, true, false>);
  }
}

// heads: this launch's 8 queue heads (the filter-only body never touches the
// per-slice done counts that follow the fused launch's heads)
void launch_filter_class(const ctsim_batch& b, const ctsim_batch::FilterClass& fc, int* heads, hipStream_t st) {
  const SliceDev* sl = b.d_slices.as<SliceDev>();
  const Item* it = b.d_itemsFc.as<Item>();
  const int* q = b.d_qFc.as<int>() + fc.qOff;
  int* c = heads;
  int* e = b.d_error.as<int>();
  const dim3 g(fc.grid), blk(256);
  switch (fc.logN) {
    case 6: hipLaunchKernelGGL(fbp_batch_filter_kernel<6>, g, blk, fc.lds, st, sl, it, q, c, e); break;
    case 7: hipLaunchKernelGGL(fbp_batch_filter_kernel<7>, g, blk, fc.lds, st, sl, it, q, c, e); break;
    case 8: hipLaunchKernelGGL(fbp_batch_filter_kernel<8>, g, blk, fc.lds, st, sl, it, q, c, e); break;
    case 9: hipLaunchKernelGGL(fbp_batch_filter_kernel<9>, g, blk, fc.lds, st, sl, it, q, c, e); break;
    case 10: hipLaunchKernelGGL(fbp_batch_filter_kernel<10>, g, blk, fc.lds, st, sl, it, q, c, e); break;
    case 11: hipLaunchKernelGGL(fbp_batch_filter_kernel<11>, g, blk, fc.lds, st, sl, it, q, c, e); break;
    case 12: hipLaunchKernelGGL(fbp_batch_filter_kernel<12>, g, blk, fc.lds, st, sl, it, q, c, e); break;
    default:
      hipLaunchKernelGGL((fbp_batch_kernel<1, false, true, false>), g, blk, fc.lds, st, sl, it, q, c, e);
  }
}

// counters: [0, 8) the fused launch's queue heads, [8, 8 + nSlices) its filter
// items done per slice, then 8 queue heads for each two-phase launch (filter
// classes, then backprojection classes), all zeroed by one reset kernel.
size_t batch_counters(const ctsim_batch& b) {
  return 8 + b.s.size() + 8 * (b.fclass.size() + b.bclass.size());
}

// The symmetric instantiation runs when some slice takes the symmetric body.
bool batch_sym_kernel(const ctsim_batch& b) {
  for (const SliceDev& d : b.host)
    if (d.sym) return true;
  return false;
}

int occupancy_grid(const void* kernel, size_t lds, int numCUs) {
  int perCU = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, 256, lds) != hipSuccess || perCU < 1)
    perCU = 1;
  int g = perCU * numCUs;
  g -= g % 8;
  return g < 8 ? 8 : g;
}

// Build the per-slice device descriptors and the item queues for the
// current shape (P, lanesY, chunk budget).  Until it succeeds the batch is
// marked dirty and holds no graph: a failed layout (a shape that does not
// fit, a failed allocation) leaves nothing to replay, and run and status
// refuse the batch until a shape is laid out.
int batch_layout(ctsim_ctx* ctx, ctsim_batch& b) {
  b.shapeDirty = true;
  if (b.graph) {
    (void)hipGraphExecDestroy(b.graph);
    b.graph = nullptr;
  }
  const int S = static_cast<int>(b.s.size());
  b.host.assign(S, SliceDev{});
  size_t maxBpLds = 0;
  std::vector<size_t> partialOff(S, 0);
  size_t partialTotal = 0;
  for (int k = 0; k < S; k++) {
    auto& sl = b.s[k];
    const BpPlan& p = sl.bp;
    SliceDev& d = b.host[k];
    d.n = p.nDet;
    d.N = sl.N;
    d.logN = sl.logN;
    d.R = (sl.N / 2 + 255) / 256;
    d.nView = p.nView;
    d.nFilterItems = (p.nView + 1) / 2;
    d.sino = b.sinos.empty() ? nullptr : b.sinos[k];
    d.w = sl.fp.preweight.empty() ? nullptr : sl.fp.d_weight.as<double>();
    if (sl.fp.method == METHOD_CONVOLUTION) {
      d.hre = sl.mul.as<float>();
      d.him = sl.mul.as<float>();
      d.tw = sl.tw.as<Cplx<float>>();
    } else {
      d.hre = sl.fp.d_hmul32.as<float>();
      d.him = sl.fp.d_hmul32.as<float>() + (sl.N / 2 + 1);
      d.tw = sl.fp.d_tw32.as<Cplx<float>>();
    }
    d.gk = p.geom == GEOM_PARALLEL ? 0 : (p.geom == GEOM_EQUIANGULAR ? 1 : 2);
    d.interp = p.interp;
    d.lastIncl = p.kind == KIND_IDIFF;
    d.deg = d.gk == 1 ? p.atanDeg : 0;
    // backprojection shape (P is the batch kernel's template parameter).
    // The (pipelined) symmetric body runs when the views come in rotated
    // copies and the windows hold at most kPipeRow cells (it stages 4 S rows
    // of kPipeRow cells) at P <= 2; in two-phase mode the symmetric classes
    // have launches of their own, so with P = 4 their slices take P = 2.
    const SymViews sv = bp_sym_views(sl.angles.data(), p.nView);
    const bool symViews = (b.sym & 7) > 0 && p.symImage && sv.S > 0;
    MixedShape m;
    m.lanesY = b.lanesY;
    m.split = 1;
    m.P = symViews && b.P == 4 && !b.fused ? 2 : b.P;
    m.W = bp_window_bound(p, 256 / m.lanesY, m.lanesY * m.P);
    d.sym = symViews && m.P <= 2 && m.W > 0 && m.W <= dev::kPipeRow;
    if (!d.sym && m.P != b.P) {  // the plain body at the batch's P after all
      m.P = b.P;
      m.W = bp_window_bound(p, 256 / m.lanesY, m.lanesY * m.P);
    }
    d.P = m.P;
    const int Wv = m.W > 0 ? m.W : p.nDet + 4;
    const size_t elem = p.interp == INTERP_CUBIC ? 16 : 8;
    // the plain body stages chunks of whole view windows within the budget;
    // the mirror body NM windows per view (when one view's fit)
    if (!d.sym && static_cast<size_t>(Wv) * elem > static_cast<size_t>(b.chunkBudget))
      return fail(CTSIM_ERR_UNSUPPORTED, "batch: a view's detector window exceeds the chunk budget");
    int chunk = static_cast<int>(b.chunkBudget / (static_cast<size_t>(Wv) * elem));
    chunk = chunk < 1 ? 1 : (chunk > dev::kMaxChunk ? dev::kMaxChunk : chunk);
    m.chunk = chunk;
    if (static_cast<size_t>(chunk) * Wv * elem > 62 * 1024)
      return fail(CTSIM_ERR_UNSUPPORTED, "batch: a slice's detector window does not fit in LDS");
    // the mirror body (two-phase mode) stages into static buffers of its
    // own, as many views per chunk as fit them (none: the plain body)
    d.mirNM = 1;
    if (!d.sym && (b.sym & 8) && !b.fused && sl.mir.NM > 1) {
      MixedShape mm = m;
      mm.chunk = dev::kMaxChunk;
      const int mc = bp_mirror_chunk(p, mm, sl.mir.NM, p.nView);
      if (mc > 0) {
        d.mirNM = sl.mir.NM;
        m.chunk = mc;
      }
    }
    const size_t symLds = static_cast<size_t>(4 * sv.S) * dev::kPipeRow * elem;
    d.symS = sv.S;
    d.symK = sv.K;
    if (d.mirNM == 1) maxBpLds = std::max(maxBpLds, d.sym ? symLds : static_cast<size_t>(chunk) * Wv * elem);
    MixedArgs& a = d.bp;
    a = MixedArgs{};
    a.nx = p.nx; a.ny = p.ny; a.nDet = p.nDet; a.nView = p.nView;
    a.lanesY = m.lanesY;
    const int tx = 256 / m.lanesY, ty = m.lanesY * m.P;
    a.tilesX = (p.nx + tx - 1) / tx;
    a.tilesY = (p.ny + ty - 1) / ty;
    a.W = Wv;
    a.chunk = m.chunk;
    a.split = 1;
    a.viewsPerGroup = p.nView;
    a.center = static_cast<float>(p.iDetCenter);
    a.F = static_cast<float>(p.F);
    a.twoF = static_cast<float>(2 * p.F);
    a.endScale = static_cast<float>(p.rotScale);
    a.posScale = d.gk == 2 ? static_cast<float>(p.sdl / p.detInc) : static_cast<float>(1.0 / p.detInc);
    if (d.deg > 0) for (int j = 0; j <= d.deg; j++) a.atan[j] = p.atanCoef[j];
    else a.atan[0] = static_cast<float>(1.0 / p.detInc);
    a.xs = p.d_xy.as<float>();
    a.ys = p.d_xy.as<float>() + p.xs.size();
    a.vc = sl.vc.as<ViewConstF32>();
    a.coef = sl.coef.ptr;
    a.image = b.images.empty() ? nullptr : b.images[k];
    a.overwrite = 1;
    a.domX = p.nx;
    a.domY = p.ny;
    if (d.mirNM > 1) {  // tiles of the mirror domain
      a.domX = (p.nx + 1) / 2;
      a.domY = d.mirNM == 4 ? (p.ny + 1) / 2 : p.ny;
      a.mapO = sl.mir.mapO;
      a.tilesX = (a.domX + tx - 1) / tx;
      a.tilesY = (a.domY + ty - 1) / ty;
    }
    if (d.sym) {
      a.half = p.nx / 2;
      a.Q = p.nView / sv.S;
      a.dirf = sv.dir;
      a.tilesX = (a.half + tx - 1) / tx;
      a.tilesY = (a.half + ty - 1) / ty;
      a.capW = 4 * sv.S * dev::kPipeRow;
      // view groups of at least 8 base views; each group's partial image
      // goes to the batch's partial buffer, added by batch_combine_kernel
      int G = std::max(1, std::min(b.sym & 7, a.Q / 8));
      a.viewsPerGroup = (a.Q + G - 1) / G;
      G = (a.Q + a.viewsPerGroup - 1) / a.viewsPerGroup;
      a.split = G;
      if (G > 1) {
        partialOff[k] = partialTotal;
        partialTotal += static_cast<size_t>(G) * p.nx * p.ny;
      }
    }
  }
  std::vector<int> combine;
  if (partialTotal > 0) {
    CTSIM_CHECK(b.d_partial.reserve(partialTotal * sizeof(float)));
    for (int k = 0; k < S; k++)
      if (b.host[k].sym && b.host[k].bp.split > 1) {
        b.host[k].bp.partial = b.d_partial.as<float>() + partialOff[k];
        combine.push_back(k);
      }
  }
  b.nCombine = static_cast<int>(combine.size());
  // Slices to XCDs, longest first, so that a slice's coefficients are made
  // and read under one L2; a slice costing more than a quarter of an XCD's
  // share is dealt item by item over all eight instead (one big slice on
  // one XCD would leave the other seven idle).
  std::vector<int> order(S);
  for (int k = 0; k < S; k++) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b.s[x].cost > b.s[y].cost; });
  double total = 0;
  for (int k = 0; k < S; k++) total += b.s[k].cost;
  std::vector<std::vector<Item>> f(8), bpq(8);
  std::vector<double> load(8, 0.0);
  auto bpItems = [&](int k, std::vector<Item>& out, int x0, int stride) {
    const int tiles = b.host[k].bp.tilesX * b.host[k].bp.tilesY;
    const int groups = b.host[k].sym ? b.host[k].bp.split : 1;
    int i = 0;
    for (int t = 0; t < tiles; t++)
      for (int g = 0; g < groups; g++, i++)
        if (stride == 1 || i % stride == x0) out.push_back({kBpItem, k, t, g});
  };
  for (int k : order) {
    if (b.s[k].cost > total / 32) {
      for (int pr = 0; pr < b.host[k].nFilterItems; pr++) f[pr % 8].push_back({kFilterItem, k, pr, 0});
      for (int x = 0; x < 8; x++) {
        bpItems(k, bpq[x], x, 8);
        load[x] += b.s[k].cost / 8;
      }
    } else {
      const int x = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
      for (int pr = 0; pr < b.host[k].nFilterItems; pr++) f[x].push_back({kFilterItem, k, pr, 0});
      bpItems(k, bpq[x], 0, 1);
      load[x] += b.s[k].cost;
    }
  }
  // the two-phase filter launches by transform length class
  std::vector<Item> itemsFc;
  std::vector<int> qFc;
  b.fclass.clear();
  auto fclassOf = [&](int k) { return dev::fft_c_length(b.host[k].logN) ? b.host[k].logN : 0; };
  for (int cls : {6, 7, 8, 9, 10, 11, 12, 0}) {
    ctsim_batch::FilterClass fc;
    fc.logN = cls;
    fc.qOff = static_cast<int>(qFc.size());
    int maxN = 0;
    const size_t first = itemsFc.size();
    for (int x = 0; x < 8; x++) {
      qFc.push_back(static_cast<int>(itemsFc.size()));
      for (const Item& it : f[x])
        if (fclassOf(it.slice) == cls) {
          itemsFc.push_back(it);
          maxN = std::max(maxN, b.host[it.slice].N);
        }
    }
    qFc.push_back(static_cast<int>(itemsFc.size()));
    if (itemsFc.size() == first) {
      qFc.resize(fc.qOff);
      continue;
    }
    fc.lds = cls ? dev::fft_c_lds(maxN)
                 : static_cast<size_t>(maxN + dev::fft_tw_entries<float>(maxN)) * sizeof(Cplx<float>);
    b.fclass.push_back(fc);
  }
  // the two-phase backprojection by class (body, geometry, interpolation,
  // polynomial degree, symmetry, P), each class's items in the per-XCD order
  // above
  // plain-body slices of several classes share one generic launch (key -1),
  // and so do mirror-body slices (key -2)
  auto plainKey = [&](int k) {
    const SliceDev& d = b.host[k];
    return d.gk + 3 * (d.interp == INTERP_CUBIC ? 1 : 0) + 6 * d.deg + 60 * d.mirNM + 900 * d.P;
  };
  std::vector<int> plainKeys, mirrorKeys;
  for (int k = 0; k < S; k++) {
    if (b.host[k].sym) continue;
    std::vector<int>& keys = b.host[k].mirNM > 1 ? mirrorKeys : plainKeys;
    if (std::find(keys.begin(), keys.end(), plainKey(k)) == keys.end()) keys.push_back(plainKey(k));
  }
  const bool genericPlain = plainKeys.size() > 1, genericMirror = mirrorKeys.size() > 1;
  auto bclassKey = [&](int k) {
    const SliceDev& d = b.host[k];
    if (!d.sym && d.mirNM > 1) return genericMirror ? -2 : plainKey(k);
    if (!d.sym) return genericPlain ? -1 : plainKey(k);
    return d.gk + 3 * (d.interp == INTERP_CUBIC ? 1 : 0) + 6 * d.deg + 60 * d.symS + 300 * d.symK +
           900 * d.P + 9000;
  };
  std::vector<int> keys;
  for (int k = 0; k < S; k++)
    if (std::find(keys.begin(), keys.end(), bclassKey(k)) == keys.end()) keys.push_back(bclassKey(k));
  std::vector<Item> itemsBc;
  std::vector<int> qBc;
  b.bclass.clear();
  for (int key : keys) {
    ctsim_batch::BpClass bc;
    bc.qOff = static_cast<int>(qBc.size());
    int rep = -1;
    for (int x = 0; x < 8; x++) {
      qBc.push_back(static_cast<int>(itemsBc.size()));
      for (const Item& it : bpq[x])
        if (bclassKey(it.slice) == key) {
          itemsBc.push_back(it);
          rep = it.slice;
        }
    }
    qBc.push_back(static_cast<int>(itemsBc.size()));
    const SliceDev& d = b.host[rep];
    const size_t elem = d.interp == INTERP_CUBIC ? 16 : 8;
    if (d.sym) {
      bc.kernel = sym_class_kernel(d.P, d.gk, d.interp, d.deg, d.symS, d.symK);
      bc.lds = static_cast<size_t>(4 * d.symS) * dev::kPipeRow * elem;
    } else {
      bc.generic = key < 0;
      bc.kernel = !bc.generic ? plain_class_kernel(d.P, d.gk, d.interp, d.deg, d.mirNM)
                  : key == -2 ? (d.P == 4   ? reinterpret_cast<const void*>(fbp_batch_mirror_kernel<4>)
                                 : d.P == 1 ? reinterpret_cast<const void*>(fbp_batch_mirror_kernel<1>)
                                            : reinterpret_cast<const void*>(fbp_batch_mirror_kernel<2>))
                  : d.P == 4  ? reinterpret_cast<const void*>(fbp_batch_bp_kernel<4>)
                  : d.P == 1  ? reinterpret_cast<const void*>(fbp_batch_bp_kernel<1>)
                              : reinterpret_cast<const void*>(fbp_batch_bp_kernel<2>);
      for (int k = 0; k < S; k++)  // the largest staging of the launch's slices
        if (bclassKey(k) == key)
          if (b.host[k].mirNM == 1)  // (the mirror body's staging is static)
            bc.lds = std::max(bc.lds, static_cast<size_t>(b.host[k].bp.chunk) * b.host[k].bp.W *
                                          (b.host[k].interp == INTERP_CUBIC ? 16 : 8));
    }
    b.bclass.push_back(bc);
  }
  std::vector<Item> itemsF, itemsAll;
  std::vector<int> qF(9), qAll(9);
  for (int x = 0; x < 8; x++) {
    qF[x] = static_cast<int>(itemsF.size());
    qAll[x] = static_cast<int>(itemsAll.size());
    itemsF.insert(itemsF.end(), f[x].begin(), f[x].end());
    itemsAll.insert(itemsAll.end(), f[x].begin(), f[x].end());
    itemsAll.insert(itemsAll.end(), bpq[x].begin(), bpq[x].end());
  }
  qF[8] = static_cast<int>(itemsF.size());
  qAll[8] = static_cast<int>(itemsAll.size());
  auto up = [](DevBuf& d, const void* src, size_t n) -> int {
    CTSIM_CHECK(d.reserve(n > 0 ? n : 16));
    if (n > 0) CTSIM_HIP(hipMemcpy(d.ptr, src, n, hipMemcpyHostToDevice));
    return CTSIM_OK;
  };
  CTSIM_CHECK(up(b.d_slices, b.host.data(), b.host.size() * sizeof(SliceDev)));
  CTSIM_CHECK(up(b.d_itemsF, itemsF.data(), itemsF.size() * sizeof(Item)));
  CTSIM_CHECK(up(b.d_itemsFc, itemsFc.data(), itemsFc.size() * sizeof(Item)));
  CTSIM_CHECK(up(b.d_qFc, qFc.data(), qFc.size() * sizeof(int)));
  if (!b.bclass.empty()) {
    CTSIM_CHECK(up(b.d_itemsBc, itemsBc.data(), itemsBc.size() * sizeof(Item)));
    CTSIM_CHECK(up(b.d_qBc, qBc.data(), qBc.size() * sizeof(int)));
  }
  CTSIM_CHECK(up(b.d_itemsAll, itemsAll.data(), itemsAll.size() * sizeof(Item)));
  CTSIM_CHECK(up(b.d_qF, qF.data(), qF.size() * sizeof(int)));
  CTSIM_CHECK(up(b.d_qAll, qAll.data(), qAll.size() * sizeof(int)));
  CTSIM_CHECK(b.d_counters.reserve(batch_counters(b) * sizeof(int)));
  if (!combine.empty()) CTSIM_CHECK(up(b.d_combine, combine.data(), combine.size() * sizeof(int)));
  CTSIM_CHECK(b.d_error.reserve(sizeof(int)));
  CTSIM_HIP(hipMemset(b.d_error.ptr, 0, sizeof(int)));
  b.ldsAll = std::max(maxBpLds, static_cast<size_t>(b.ldsFilter));
  const bool symKernel = batch_sym_kernel(b);
  auto pick = [&](auto k4, auto k1s, auto k1, auto k2s, auto k2) -> const void* {
    return b.P == 4 ? reinterpret_cast<const void*>(k4)
           : b.P == 1 ? (symKernel ? reinterpret_cast<const void*>(k1s) : reinterpret_cast<const void*>(k1))
                      : (symKernel ? reinterpret_cast<const void*>(k2s) : reinterpret_cast<const void*>(k2));
  };
  const void* kAll = pick(fbp_batch_kernel<4, false, true, true>, fbp_batch_kernel<1, true, true, true>,
                          fbp_batch_kernel<1, false, true, true>, fbp_batch_kernel<2, true, true, true>,
                          fbp_batch_kernel<2, false, true, true>);
  const void* kF = reinterpret_cast<const void*>(fbp_batch_kernel<1, false, true, false>);
  b.gridF = occupancy_grid(kF, b.ldsFilter, ctx->numCUs);
  for (auto& fc : b.fclass) fc.grid = occupancy_grid(filter_kernel_of(fc.logN), fc.lds, ctx->numCUs);
  for (auto& bc : b.bclass) bc.grid = occupancy_grid(bc.kernel, bc.lds, ctx->numCUs);
  b.gridAll = occupancy_grid(kAll, b.ldsAll, ctx->numCUs);
  b.shapeDirty = false;
  return CTSIM_OK;
}

template <int P, bool SYM>
void launch_batch(ctsim_batch& b, hipStream_t st) {
  const int S = static_cast<int>(b.s.size());
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(256), 0, st, b.d_counters.as<int>(),
                     static_cast<int>(batch_counters(b)), b.d_error.as<int>());
  if (b.fused) {
    hipLaunchKernelGGL((fbp_batch_kernel<P, SYM, true, true>), dim3(b.gridAll), dim3(256), b.ldsAll, st,
                       b.d_slices.as<SliceDev>(), b.d_itemsAll.as<Item>(), b.d_qAll.as<int>(),
                       b.d_counters.as<int>(), b.d_error.as<int>());
  } else {
    // one filter launch per transform length class, then one backprojection
    // launch per class, each with queue heads of its own
    int* heads = b.d_counters.as<int>() + 8 + S;
    for (size_t c = 0; c < b.fclass.size(); c++, heads += 8) launch_filter_class(b, b.fclass[c], heads, st);
    for (const auto& bc : b.bclass) {
      const SliceDev* sl = b.d_slices.as<SliceDev>();
      const Item* it = b.d_itemsBc.as<Item>();
      const int* q = b.d_qBc.as<int>() + bc.qOff;
      int* c = heads;
      int* e = b.d_error.as<int>();
      void* args[] = {&sl, &it, &q, &c, &e};
      (void)hipLaunchKernel(bc.kernel, dim3(bc.grid), dim3(256), args, bc.lds, st);
      heads += 8;
    }
  }
  if (b.nCombine > 0)
    hipLaunchKernelGGL(batch_combine_kernel, dim3(256, b.nCombine), dim3(256), 0, st,
                       b.d_slices.as<SliceDev>(), b.d_combine.as<int>());
}

int batch_enqueue(ctsim_ctx* ctx, ctsim_batch& b) {
  // capture once per layout, replay afterwards
  if (!b.graph) {
    hipGraph_t g = nullptr;
    CTSIM_HIP(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
    const bool sym = batch_sym_kernel(b);
    if (b.P == 4) launch_batch<4, false>(b, b.stream);
    else if (b.P == 1) sym ? launch_batch<1, true>(b, b.stream) : launch_batch<1, false>(b, b.stream);
    else sym ? launch_batch<2, true>(b, b.stream) : launch_batch<2, false>(b, b.stream);
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(b.stream, &g);
    if (le != hipSuccess) return hip_fail(le, "batch launch", __FILE__, __LINE__);
    if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture", __FILE__, __LINE__);
    const hipError_t ie = hipGraphInstantiate(&b.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
      b.graph = nullptr;
      return hip_fail(ie, "hipGraphInstantiate", __FILE__, __LINE__);
    }
  }
  CTSIM_HIP(hipEventRecord(b.evIn, ctx->stream));
  CTSIM_HIP(hipStreamWaitEvent(b.stream, b.evIn, 0));
  ctx->prof_begin_on(b.stream);
  CTSIM_HIP(hipGraphLaunch(b.graph, b.stream));
  ctx->prof_end_on(b.stream);
  CTSIM_HIP(hipEventRecord(b.evOut, b.stream));
  CTSIM_HIP(hipStreamWaitEvent(ctx->stream, b.evOut, 0));
  return CTSIM_OK;
}

}  // namespace
}  // namespace ctsim

extern "C" {

int ctsim_batch_create(ctsim_ctx* ctx, ctsim_batch** out, const ctsim_slice_desc* d, int nSlices) {
  if (!ctx || !out || (!d && nSlices > 0)) return fail(CTSIM_ERR_INVALID, "ctsim_batch_create: NULL argument");
  if (nSlices < 1) return fail(CTSIM_ERR_INVALID, "ctsim_batch_create: no slices");
  *out = nullptr;
  CTSIM_CHECK(ctx->activate());
  auto* b = new ctsim_batch();
  b->s.resize(nSlices);
  int rc = CTSIM_OK;
  int maxN = 0;
  for (int k = 0; k < nSlices && rc == CTSIM_OK; k++) {
    const ctsim_slice_desc& sd = d[k];
    auto& sl = b->s[k];
    if (!sd.angles) { rc = fail(CTSIM_ERR_INVALID, "ctsim_batch_create: slice without view angles"); break; }
    rc = filter_plan_init(sl.fp, sd.filter, sd.method, 1.0 / sd.detInc, sd.detInc, sd.nDet, sd.filterParam,
                          "spatial", sd.generation, sd.zeropad, 1, sd.geometry, sd.focalLength,
                          sd.sourceDetectorLength);
    if (rc == CTSIM_OK) rc = filter_upload(sl.fp);
    if (rc == CTSIM_OK)
      rc = bp_plan_init(sl.bp, sd.geometry, sd.nDet, sd.detInc, sd.nView, sd.viewDiameter,
                        sd.focalLength, sd.sourceDetectorLength, sd.backproj, sd.interp, 1, sd.nx,
                        sd.ny, nullptr);
    if (rc == CTSIM_OK) rc = bp_upload_tables(sl.bp);
    if (rc != CTSIM_OK) break;
    if (sl.bp.interp == INTERP_NEAREST || sl.bp.kind == KIND_DIFF) {
      rc = fail(CTSIM_ERR_UNSUPPORTED, "batch: nearest interpolation and Diff take the per-slice FP64 index path");
      break;
    }
    const int n = sd.nDet;
    if (sl.fp.method == METHOD_CONVOLUTION) {
      // FFT convolution: circular length >= 2n-1 holds every lag of the
      // reference's direct sum (procsignal.cpp:659-673) without wrap-around;
      // the multipliers are the real part of the kernel's transform (the
      // kernel is even up to the rounding of its sample positions)
      int N = 1;
      while (N < 2 * n - 1) N <<= 1;
      sl.N = N;
      std::vector<std::complex<double>> g(N, 0.0);
      for (int lag = -(n - 1); lag <= n - 1; lag++)
        g[(lag + N) % N] = sl.fp.inc * sl.fp.filter[lag + n - 1];
      fft_host(g);
      std::vector<float> mul(N / 2 + 1);
      for (int kk = 0; kk <= N / 2; kk++) mul[kk] = static_cast<float>(g[kk].real() / N);
      std::vector<float> tw(N);
      for (int kk = 0; kk < N / 2; kk++) {
        const double a = -2.0 * kPi * kk / N;
        tw[2 * kk] = static_cast<float>(cos(a));
        tw[2 * kk + 1] = static_cast<float>(sin(a));
      }
      dev::fft_append_pass_twiddles(tw, N);
      if ((rc = sl.mul.reserve(mul.size() * sizeof(float))) != CTSIM_OK) break;
      if ((rc = sl.tw.reserve(tw.size() * sizeof(float))) != CTSIM_OK) break;
      if (hipMemcpy(sl.mul.ptr, mul.data(), mul.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(sl.tw.ptr, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(CTSIM_ERR_HIP, "batch: table upload failed");
        break;
      }
    } else {
      sl.N = sl.fp.nfp;
      if ((sl.N & (sl.N - 1)) != 0) {
        rc = fail(CTSIM_ERR_UNSUPPORTED, "batch: frequency filtering needs a power-of-two transform (zeropad >= 1)");
        break;
      }
    }
    while ((1 << sl.logN) < sl.N) sl.logN++;
    if (sl.N > 4096) { rc = fail(CTSIM_ERR_UNSUPPORTED, "batch: transform longer than 4096"); break; }
    maxN = std::max(maxN, sl.N);
    sl.angles.assign(sd.angles, sd.angles + sd.nView);
    std::vector<ViewConstF32> vc(sd.nView);
    sl.mir = bp_view_consts_f32(sl.bp, sd.angles, sd.nView, vc.data());
    if ((rc = sl.vc.reserve(vc.size() * sizeof(ViewConstF32))) != CTSIM_OK) break;
    if (hipMemcpy(sl.vc.ptr, vc.data(), vc.size() * sizeof(ViewConstF32), hipMemcpyHostToDevice) != hipSuccess) {
      rc = fail(CTSIM_ERR_HIP, "batch: view constant upload failed");
      break;
    }
    const size_t elem = sl.bp.interp == INTERP_CUBIC ? 16 : 8;
    if ((rc = sl.coef.reserve(static_cast<size_t>(sd.nView) * sd.nDet * elem)) != CTSIM_OK) break;
    const double geomCost = sl.bp.geom == GEOM_PARALLEL ? 1.0 : (sl.bp.geom == GEOM_EQUIANGULAR ? 3.0 : 2.0);
    const double pvu = static_cast<double>(sd.nx) * sd.ny * sd.nView;
    sl.cost = pvu * geomCost * (sl.bp.interp == INTERP_CUBIC ? 1.5 : 1.0) +
              40.0 * sd.nView * sl.N * sl.logN;
    b->pvu += pvu;
  }
  if (rc == CTSIM_OK) {
    // the fused launch's filter items: compile-time lengths hold N twiddles
    size_t ldsF = 0;
    for (const auto& sl : b->s)
      ldsF = std::max(ldsF, dev::fft_c_length(sl.logN)
                                ? dev::fft_c_lds(sl.N)
                                : static_cast<size_t>(sl.N + dev::fft_tw_entries<float>(sl.N)) * sizeof(Cplx<float>));
    b->ldsFilter = static_cast<int>(ldsF);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->evIn, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->evOut, hipEventDisableTiming) != hipSuccess)
      rc = fail(CTSIM_ERR_HIP, "batch: stream or event creation failed");
  }
  if (rc == CTSIM_OK) rc = batch_layout(ctx, *b);
  if (rc == CTSIM_ERR_UNSUPPORTED && b->chunkBudget < 62 * 1024) {
    // a detector window wider than the default chunk budget: the largest
    b->chunkBudget = 62 * 1024;
    rc = batch_layout(ctx, *b);
  }
  if (rc != CTSIM_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return CTSIM_OK;
}

void ctsim_batch_destroy(ctsim_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  delete b;
}

int ctsim_batch_set_shape(ctsim_ctx* ctx, ctsim_batch* b, int P, int lanesY, int chunkBytes, int fused,
                          int sym) {
  if (!ctx || !b) return fail(CTSIM_ERR_INVALID, "ctsim_batch_set_shape: NULL argument");
  if (!(P == 1 || P == 2 || P == 4) || !(lanesY == 8 || lanesY == 16 || lanesY == 32) ||
      chunkBytes < 1024 || chunkBytes > 62 * 1024 ||
      !((sym & ~8) == 0 || (sym & ~8) == 1 || (sym & ~8) == 2 || (sym & ~8) == 4) || sym < 0 || sym > 12)
    return fail(CTSIM_ERR_INVALID, "ctsim_batch_set_shape: shape out of range");
  CTSIM_CHECK(ctx->activate());
  CTSIM_HIP(hipStreamSynchronize(b->stream));
  b->P = P;
  b->lanesY = lanesY;
  b->chunkBudget = chunkBytes;
  b->fused = fused != 0;
  b->sym = sym;
  return batch_layout(ctx, *b);
}

int ctsim_batch_run(ctsim_ctx* ctx, ctsim_batch* b, const float* const* d_sinos, float* const* d_images) {
  if (!ctx || !b || !d_sinos || !d_images) return fail(CTSIM_ERR_INVALID, "ctsim_batch_run: NULL argument");
  const int S = static_cast<int>(b->s.size());
  for (int k = 0; k < S; k++)
    if (!d_sinos[k] || !d_images[k]) return fail(CTSIM_ERR_INVALID, "ctsim_batch_run: NULL slice pointer");
  CTSIM_CHECK(ctx->activate());
  if (b->shapeDirty)
    return fail(CTSIM_ERR_INVALID, "ctsim_batch_run: the batch's last layout failed; set a supported shape");
  const bool same = static_cast<int>(b->sinos.size()) == S &&
                    std::equal(b->sinos.begin(), b->sinos.end(), d_sinos) &&
                    std::equal(b->images.begin(), b->images.end(), d_images);
  if (!same) {
    CTSIM_HIP(hipStreamSynchronize(b->stream));
    b->sinos.assign(d_sinos, d_sinos + S);
    b->images.assign(d_images, d_images + S);
    for (int k = 0; k < S; k++) {
      b->host[k].sino = d_sinos[k];
      b->host[k].bp.image = d_images[k];
    }
    CTSIM_HIP(hipMemcpy(b->d_slices.ptr, b->host.data(), S * sizeof(SliceDev), hipMemcpyHostToDevice));
  }
  return batch_enqueue(ctx, *b);
}

int ctsim_batch_status(ctsim_ctx* ctx, ctsim_batch* b) {
  if (!ctx || !b) return fail(CTSIM_ERR_INVALID, "ctsim_batch_status: NULL argument");
  CTSIM_CHECK(ctx->activate());
  if (b->shapeDirty)
    return fail(CTSIM_ERR_INVALID, "ctsim_batch_status: the batch's last layout failed; set a supported shape");
  CTSIM_HIP(hipStreamSynchronize(b->stream));
  int flag = 0;
  CTSIM_HIP(hipMemcpy(&flag, b->d_error.ptr, sizeof(int), hipMemcpyDeviceToHost));
  if (flag) return fail(CTSIM_ERR_HIP, "batch: a backprojection item timed out waiting for its filter items");
  return CTSIM_OK;
}

int ctsim_batch_autotune(ctsim_ctx* ctx, ctsim_batch* b, const float* const* d_sinos,
                         float* const* d_images, int* shape5, double* bestMs) {
  if (!ctx || !b) return fail(CTSIM_ERR_INVALID, "ctsim_batch_autotune: NULL argument");
  CTSIM_CHECK(ctx->activate());
  hipEvent_t e0, e1;
  CTSIM_HIP(hipEventCreate(&e0));
  CTSIM_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int bP = b->P, bL = b->lanesY, bC = b->chunkBudget, bF = b->fused, bS = b->sym;
  int rc = CTSIM_OK;
  // P = 1 only with the symmetric body; P = 4 with it only in two-phase
  // mode (its symmetric slices then take P = 2 in launches of their own)
  // (the mirror body, sym + 8, only in two-phase mode)
  for (int P : {1, 2, 4})
    for (int sym : {1, 4, 0, 9, 12, 8})
    for (int lanesY : {8, 16, 32})
      for (int chunk : {16 * 1024, 32 * 1024})
        for (int fused : {1, 0}) {
          if (rc != CTSIM_OK) break;
          if ((P == 1 && !(sym & 7)) || (P == 4 && (sym & 7) && fused) || ((sym & 8) && fused)) continue;
          if ((rc = ctsim_batch_set_shape(ctx, b, P, lanesY, chunk, fused, sym)) != CTSIM_OK) {
            if (rc == CTSIM_ERR_UNSUPPORTED) { rc = CTSIM_OK; continue; }
            break;
          }
          // a warm-up run, then two timed runs; a shape whose runs time out
          // waiting for filter items (the status check) is skipped
          float tot = 0;
          bool timedOut = false;
          for (int rep = 0; rep < 3 && rc == CTSIM_OK && !timedOut; rep++) {
            if (rep > 0) (void)hipEventRecord(e0, ctx->stream);
            if ((rc = ctsim_batch_run(ctx, b, d_sinos, d_images)) != CTSIM_OK) break;
            if (rep > 0) {
              (void)hipEventRecord(e1, ctx->stream);
              (void)hipEventSynchronize(e1);
            }
            if (ctsim_batch_status(ctx, b) != CTSIM_OK) {
              timedOut = true;
              break;
            }
            if (rep == 0) continue;
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            tot += ms;
          }
          if (rc == CTSIM_OK && !timedOut && tot / 2 < best) {
            best = tot / 2;
            bP = P; bL = lanesY; bC = chunk; bF = fused; bS = sym;
          }
        }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != CTSIM_OK) return rc;
  CTSIM_CHECK(ctsim_batch_set_shape(ctx, b, bP, bL, bC, bF, bS));
  if (shape5) { shape5[0] = bP; shape5[1] = bL; shape5[2] = bC; shape5[3] = bF; shape5[4] = bS; }
  if (bestMs) *bestMs = best;
  return CTSIM_OK;
}

int ctsim_batch_info(const ctsim_batch* b, double* out4) {
  if (!b || !out4) return fail(CTSIM_ERR_INVALID, "ctsim_batch_info: NULL argument");
  out4[0] = static_cast<double>(b->s.size());
  out4[1] = b->pvu;
  out4[2] = b->fused ? 1 : 0;
  out4[3] = b->gridAll;
  return CTSIM_OK;
}

}  // extern "C"
