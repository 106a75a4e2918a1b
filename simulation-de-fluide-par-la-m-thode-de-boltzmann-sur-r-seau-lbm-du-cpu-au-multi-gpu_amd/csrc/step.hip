// Fused D2Q9 time step for gfx950 (float and double) and its read-back kernels.
//
// The hot path of the reference is step_kernel<kPull=true, kNoModulo=true, kOpen=false>
// (src/v3_cuda_opt/solver_v3.cu:45-113): for every fluid cell, pull the nine populations from
// the neighbours (halfway bounce-back from solid ones), compute the moments, relax with BGK plus
// the Guo source, and store to the cell's own slot. This file restates that update for CDNA4:
//
//   * A workgroup owns a tile of block_x consecutive columns (a multiple of the 64-wide
//     wavefront) by rows_per_tile rows; each thread owns one column of the tile and walks its
//     rows. Row indices are wave-uniform: the row offsets live in SGPRs (buffer soffset) and each
//     lane only carries three column byte offsets (x - 1, x, x + 1, wrapped once per thread with
//     a compare, never a modulo).
//   * Workgroups run in the hardware's round-robin XCD order by default; xcd_bands remaps them
//     so that each of the 8 XCDs walks a contiguous band of tiles (measured no faster here).
//   * Solid cells, bounce-back links and open-x edges are handled only in "dirty" 64-column
//     segments (one scalar flag per wave and row); clean segments, the bulk of a channel, read
//     nothing but the nine populations: 72 B per cell update in fp32, the algorithmic minimum.
//   * Per-solver parameters travel as kernel arguments (no __constant__ globals).
//
// Floating point: -ffp-contract=off and IEEE division, so float and double steps are
// bit-identical to the reference's v0.
#include "step_body.hpp"
#include "kernels.hpp"

namespace lbmhip {
namespace {

template <class T, bool kOpen, bool kField, int kStoreAux, bool kOne = false>
__global__ void __launch_bounds__(1024) step_kernel(const StepParams<T> p) {
  const unsigned tile = tile_in_order(p.xcd_bands, blockIdx.x, gridDim.x, static_cast<unsigned>(p.tiles_x));
  const int tx = static_cast<int>(tile % static_cast<unsigned>(p.tiles_x));
  const int ty = static_cast<int>(tile / static_cast<unsigned>(p.tiles_x));
  const int x = tx * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
  const int y0 = p.row0 + ty * p.rows_per_tile;
  const int y1 = min(y0 + p.rows_per_tile, p.row0 + p.nrows);
  const DomainView<T> d{p.g,  p.gn, p.links, p.segflags, p.force,  p.N,  p.nx,
                        p.ny, p.nseg, p.wrap_y, p.omega, p.fp,    p.fx, p.fy};
  using PS = typename std::conditional<kOne, PlanesOne, Planes>::type;
  update_column<T, kOpen, kField, kStoreAux, PS>(d, x, y0, y1, p.eq_in);
}

// ---- read-back kernels: one thread per cell, gather with the per-cell link code -------------

template <class T>
__global__ void gather_kernel(const GatherParams<T> p, T* __restrict__ out) {
  const size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= static_cast<size_t>(p.nx) * p.ny) return;
  const int x = static_cast<int>(c % p.nx), y = static_cast<int>(c / p.nx);
  T v[kQ];
  if (!gather_cell<T>(p.g, p.links, p.N, p.nx, p.ny, p.eq_in, x, y, v))
#pragma unroll
    for (int i = 0; i < kQ; ++i) v[i] = T(0);
#pragma unroll
  for (int i = 0; i < kQ; ++i) out[c * kQ + i] = v[i];
}

template <class T>
__global__ void gather_one_kernel(const GatherParams<T> p, int x, int y, T* __restrict__ out) {
  T v[kQ];
  if (!gather_cell<T>(p.g, p.links, p.N, p.nx, p.ny, p.eq_in, x, y, v))
#pragma unroll
    for (int i = 0; i < kQ; ++i) v[i] = T(0);
#pragma unroll
  for (int i = 0; i < kQ; ++i) out[i] = v[i];
}

template <class T>
__global__ void macro_kernel(const GatherParams<T> p, T* __restrict__ rho, T* __restrict__ ux,
                             T* __restrict__ uy) {
  const size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= static_cast<size_t>(p.nx) * p.ny) return;
  const int x = static_cast<int>(c % p.nx), y = static_cast<int>(c / p.nx);
  T v[kQ];
  T r = T(0), u = T(0), w = T(0);
  if (gather_cell<T>(p.g, p.links, p.N, p.nx, p.ny, p.eq_in, x, y, v)) {
    const T fx = p.force ? p.force[c] : p.fx;
    const T fy = p.force ? p.force[p.N + c] : p.fy;
    moments(v, fx, fy, r, u, w);
  }
  rho[c] = r;
  ux[c] = u;
  uy[c] = w;
}

template <class T>
__global__ void __launch_bounds__(kReduceThreads) reduce_kernel(const GatherParams<T> p,
                                                                double* __restrict__ partials) {
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  const size_t cells = static_cast<size_t>(p.nx) * p.ny;
  for (size_t c = static_cast<size_t>(blockIdx.x) * kReduceThreads + threadIdx.x; c < cells;
       c += static_cast<size_t>(gridDim.x) * kReduceThreads) {
    const int x = static_cast<int>(c % p.nx), y = static_cast<int>(c / p.nx);
    T v[kQ];
    if (!gather_cell<T>(p.g, p.links, p.N, p.nx, p.ny, p.eq_in, x, y, v)) continue;
    s[0] += 1.0;
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      const double d = static_cast<double>(v[i]);
      s[1] += d;
      s[2] += kCx[i] * d;
      s[3] += kCy[i] * d;
    }
  }
  block_sum4(s);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k) partials[blockIdx.x * 4 + k] = s[k];
}

// Grid-stride copy, the counterpart of the reference's ceiling kernel
// (bench/gpu_bandwidth.cu:17-22); kVec floats per access (1 = the reference's form, 4 = one
// 16-byte access per lane), optionally with non-temporal hints.
template <int kVec, bool kNt>
__global__ void copy_kernel(const float* __restrict__ a, float* __restrict__ b, size_t n) {
  using V = float __attribute__((ext_vector_type(kVec)));
  const V* av = reinterpret_cast<const V*>(a);
  V* bv = reinterpret_cast<V*>(b);
  const size_t nv = n / kVec;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < nv;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    if constexpr (kNt)
      __builtin_nontemporal_store(__builtin_nontemporal_load(av + i), bv + i);
    else
      bv[i] = av[i];
  }
}

template <bool kNt>
__global__ void __launch_bounds__(256) copy9_kernel(const float* __restrict__ a, float* __restrict__ b,
                                                    size_t n9) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n9) return;
  float v[kQ];
#pragma unroll
  for (int k = 0; k < kQ; ++k) v[k] = kNt ? __builtin_nontemporal_load(a + k * n9 + i) : a[k * n9 + i];
#pragma unroll
  for (int k = 0; k < kQ; ++k) {
    if constexpr (kNt)
      __builtin_nontemporal_store(v[k], b + k * n9 + i);
    else
      b[k * n9 + i] = v[k];
  }
}

// Ghost rows of a strip whose ring is closed on itself, in one launch: every entry of the
// message the strip sends up lands where a message from the strip below would, and every
// entry sent down where one from above would (halo_send / halo_recv): the contents a halo
// exchange with the neighbouring strips delivers (src/v4_mpi_cuda/engine.hpp:17-20 for one
// ghost row). blockIdx.y = direction * entries + entry; ghost rows are never read by a copy.
template <class T>
__global__ void __launch_bounds__(256) halo_ring_kernel(T* __restrict__ g, size_t N, int nx,
                                                        int nyl, int ghost) {
  const int x = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
  if (x >= nx) return;
  const int n = halo_entries(ghost);
  const int up = static_cast<int>(blockIdx.y) < n ? 1 : 0;
  const int k = static_cast<int>(blockIdx.y) - (up ? 0 : n);
  int plane = 0, from = 0, to = 0;
  halo_send(ghost, nyl, up, k, plane, from);
  halo_recv(ghost, nyl, up, k, plane, to);  // sent up = received from below
  g[plane * N + static_cast<size_t>(to) * nx + x] = g[plane * N + static_cast<size_t>(from) * nx + x];
}

}  // namespace

template <class T>
hipError_t launch_halo_ring(T* g, size_t N, int nx, int nyl, int ghost, hipStream_t stream) {
  if ((ghost != 1 && ghost != 2) || nyl < ghost || nx < 1) return hipErrorInvalidValue;
  const dim3 grid(static_cast<unsigned>((nx + 255) / 256), 2 * halo_entries(ghost));
  halo_ring_kernel<T><<<grid, 256, 0, stream>>>(g, N, nx, nyl, ghost);
  return hipGetLastError();
}
template hipError_t launch_halo_ring<float>(float*, size_t, int, int, int, hipStream_t);
template hipError_t launch_halo_ring<double>(double*, size_t, int, int, int, hipStream_t);

__global__ void __launch_bounds__(kReduceThreads)
    reduce_final_kernel(const double* __restrict__ partials, int nblocks, double* __restrict__ out) {
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblocks; b += kReduceThreads)
#pragma unroll
    for (int k = 0; k < 4; ++k) s[k] += partials[b * 4 + k];
  block_sum4(s);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = s[k];
}

template <class T>
hipError_t launch_step(const StepParams<T>& p, bool open_x, int block_x, hipStream_t stream) {
  const int tiles_y = (p.nrows + p.rows_per_tile - 1) / p.rows_per_tile;
  const unsigned grid = static_cast<unsigned>(p.tiles_x) * static_cast<unsigned>(tiles_y);
  if (grid == 0) return hipSuccess;
  const bool field = p.force != nullptr;
  if (open_x) {
    if (field)
      step_kernel<T, true, true, kAuxNone><<<grid, block_x, 0, stream>>>(p);
    else
      step_kernel<T, true, false, kAuxNone><<<grid, block_x, 0, stream>>>(p);
  } else if (field) {
    step_kernel<T, false, true, kAuxNone><<<grid, block_x, 0, stream>>>(p);
  } else {
    // the hot path: one resource for all planes when they span < 4 GiB (fewer SGPRs)
    const bool one = 9 * p.N * sizeof(T) < (size_t(1) << 32) && p.one_rsrc;
    if (p.store_nt) {
      if (one)
        step_kernel<T, false, false, kAuxNt, true><<<grid, block_x, 0, stream>>>(p);
      else
        step_kernel<T, false, false, kAuxNt, false><<<grid, block_x, 0, stream>>>(p);
    } else if (one) {
      step_kernel<T, false, false, kAuxNone, true><<<grid, block_x, 0, stream>>>(p);
    } else {
      step_kernel<T, false, false, kAuxNone, false><<<grid, block_x, 0, stream>>>(p);
    }
  }
  return hipGetLastError();
}
template hipError_t launch_step<float>(const StepParams<float>&, bool, int, hipStream_t);
template hipError_t launch_step<double>(const StepParams<double>&, bool, int, hipStream_t);

template <class T>
hipError_t launch_gather(const GatherParams<T>& p, T* out, hipStream_t stream) {
  const size_t cells = static_cast<size_t>(p.nx) * p.ny;
  gather_kernel<T><<<static_cast<unsigned>((cells + 255) / 256), 256, 0, stream>>>(p, out);
  return hipGetLastError();
}
template hipError_t launch_gather<float>(const GatherParams<float>&, float*, hipStream_t);
template hipError_t launch_gather<double>(const GatherParams<double>&, double*, hipStream_t);

template <class T>
hipError_t launch_gather_one(const GatherParams<T>& p, int x, int y, T* out9, hipStream_t stream) {
  gather_one_kernel<T><<<1, 1, 0, stream>>>(p, x, y, out9);
  return hipGetLastError();
}
template hipError_t launch_gather_one<float>(const GatherParams<float>&, int, int, float*,
                                             hipStream_t);
template hipError_t launch_gather_one<double>(const GatherParams<double>&, int, int, double*,
                                              hipStream_t);

template <class T>
hipError_t launch_macro(const GatherParams<T>& p, T* rho, T* ux, T* uy, hipStream_t stream) {
  const size_t cells = static_cast<size_t>(p.nx) * p.ny;
  macro_kernel<T><<<static_cast<unsigned>((cells + 255) / 256), 256, 0, stream>>>(p, rho, ux, uy);
  return hipGetLastError();
}
template hipError_t launch_macro<float>(const GatherParams<float>&, float*, float*, float*,
                                        hipStream_t);
template hipError_t launch_macro<double>(const GatherParams<double>&, double*, double*, double*,
                                         hipStream_t);

template <class T>
hipError_t launch_reduce(const GatherParams<T>& p, double* partials, double* out,
                         hipStream_t stream) {
  reduce_kernel<T><<<kReduceBlocks, kReduceThreads, 0, stream>>>(p, partials);
  reduce_final_kernel<<<1, kReduceThreads, 0, stream>>>(partials, kReduceBlocks, out);
  return hipGetLastError();
}
template hipError_t launch_reduce<float>(const GatherParams<float>&, double*, double*,
                                         hipStream_t);
template hipError_t launch_reduce<double>(const GatherParams<double>&, double*, double*,
                                          hipStream_t);

hipError_t launch_copy9(const float* a, float* b, size_t n9, int nt, hipStream_t stream) {
  const unsigned grid = static_cast<unsigned>((n9 + 255) / 256);
  if (nt)
    copy9_kernel<true><<<grid, 256, 0, stream>>>(a, b, n9);
  else
    copy9_kernel<false><<<grid, 256, 0, stream>>>(a, b, n9);
  return hipGetLastError();
}

hipError_t launch_copy(const float* a, float* b, size_t n, int grid, int block, int variant,
                       hipStream_t stream) {
  switch (variant) {
    case 0: copy_kernel<1, false><<<grid, block, 0, stream>>>(a, b, n); break;
    case 1: copy_kernel<4, false><<<grid, block, 0, stream>>>(a, b, n); break;
    default: copy_kernel<4, true><<<grid, block, 0, stream>>>(a, b, n); break;
  }
  return hipGetLastError();
}

}  // namespace lbmhip
