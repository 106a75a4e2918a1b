// Device-side building blocks shared by the step kernels: XCD-aware tile order, buffer
// addressing of the population planes, and the pull gather with its boundary corrections.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "d2q9.hpp"

namespace lbmhip {

constexpr unsigned kXcds = 8;  // gfx950: 8 XCDs, workgroups dispatched round-robin

// Tile index of workgroup b out of nb such that XCD k (= b % 8) gets a contiguous range of
// tiles: neighbouring tiles, which share the cache lines of the +-1 column shift of the pull,
// then run on the same XCD and meet in its L2.
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb) {
  const unsigned q = nb / kXcds, r = nb % kXcds;
  const unsigned xcd = b % kXcds, k = b / kXcds;
  return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + k;
}

// Tile of workgroup b when every XCD takes whole rows of tiles: XCD k = b % 8 walks tile rows
// k, k + 8, k + 16, ... So all tiles of a row, whose edge cache lines the +-1 column shift of the
// pull shares (the periodic ends included), meet in one L2, while the eight XCDs stay on eight
// adjacent rows (the DRAM pages of the round-robin order). Needs a multiple of 8 tile rows.
__device__ __forceinline__ unsigned xcd_row_tile(unsigned b, unsigned tiles_x) {
  const unsigned xcd = b % kXcds, k = b / kXcds;
  const unsigned group = k / tiles_x, tx = k - group * tiles_x;
  return (group * kXcds + xcd) * tiles_x + tx;
}

// Tile of workgroup b for a tile order: 0 hardware round-robin, 1 XCD bands, 2 XCD rows.
__device__ __forceinline__ unsigned tile_in_order(int order, unsigned b, unsigned nb,
                                                  unsigned tiles_x) {
  return order == 1 ? xcd_tile(b, nb) : (order == 2 ? xcd_row_tile(b, tiles_x) : b);
}

// ---- buffer addressing ----------------------------------------------------------------------
// Each population plane gets its own buffer resource (4 SGPRs). A load is
//   buffer_load v, v_col, s[plane], s_row offen
// with the wave-uniform row byte offset in an SGPR and the lane's column byte offset in one
// VGPR: no per-lane 64-bit address arithmetic at all (a plain pointer walk over rows is turned
// by the compiler into per-lane 64-bit pointers). A plane is < 4 GiB, so 32-bit offsets hold.
using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;  // gfx9-family raw buffer
constexpr int kAuxNone = 0;
constexpr int kAuxNt = 2;  // non-temporal (streaming) hint

__device__ __forceinline__ Rsrc make_rsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, -1, kRsrcWord3);
}

typedef int v2i32 __attribute__((ext_vector_type(2)));

// Storage traits: S is the stored element, C the compute type returned by load().
template <class S>
struct Stored;
template <>
struct Stored<float> {
  using C = float;
  static __device__ __forceinline__ float load(Rsrc r, unsigned vo, unsigned so) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, kAuxNone));
  }
  template <int kAux = kAuxNone>
  static __device__ __forceinline__ void store(Rsrc r, unsigned vo, unsigned so, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, vo, so, kAux);
  }
};
template <>
struct Stored<double> {
  using C = double;
  static __device__ __forceinline__ double load(Rsrc r, unsigned vo, unsigned so) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, kAuxNone));
  }
  template <int kAux = kAuxNone>
  static __device__ __forceinline__ void store(Rsrc r, unsigned vo, unsigned so, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2i32, v), r, vo, so, kAux);
  }
};
// FP16 storage: IEEE binary16 bits of the scaled population; compute in fp32 scaled units.
template <>
struct Stored<uint16_t> {
  using C = float;
  static __device__ __forceinline__ float load(Rsrc r, unsigned vo, unsigned so) {
    const unsigned short h = __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, kAuxNone);
    return static_cast<float>(__builtin_bit_cast(_Float16, h));
  }
  // the stored value itself, as a binary16 number (no conversion)
  static __device__ __forceinline__ _Float16 load_h(Rsrc r, unsigned vo, unsigned so) {
    return __builtin_bit_cast(_Float16, __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, kAuxNone));
  }
  template <int kAux = kAuxNone>
  static __device__ __forceinline__ void store_bits(Rsrc r, unsigned vo, unsigned so,
                                                    unsigned short h) {
    __builtin_amdgcn_raw_buffer_store_b16(h, r, vo, so, kAux);
  }
};

// A stored value as the two-step walk carries it: float / double as computed, binary16
// storage unconverted (the FP16 cell update converts where it reads).
template <class S>
struct WalkValue {
  using type = typename Stored<S>::C;
};
template <>
struct WalkValue<uint16_t> {
  using type = _Float16;
};

// Nine plane resources of one population buffer (any size: each plane < 4 GiB).
struct Planes {
  Rsrc r[kQ];
  template <class S>
  __device__ __forceinline__ void init(const S* base, size_t N) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) r[i] = make_rsrc(base + i * N);
  }
  template <class S>
  __device__ __forceinline__ typename Stored<S>::C load(int i, unsigned vo, unsigned so) const {
    return Stored<S>::load(r[i], vo, so);
  }
  template <class S>
  __device__ __forceinline__ typename WalkValue<S>::type load_v(int i, unsigned vo, unsigned so) const {
    if constexpr (std::is_same_v<S, uint16_t>)
      return Stored<uint16_t>::load_h(r[i], vo, so);
    else
      return Stored<S>::load(r[i], vo, so);
  }
  template <class S, int kAux>
  __device__ __forceinline__ void store(int i, unsigned vo, unsigned so, typename Stored<S>::C v) const {
    Stored<S>::template store<kAux>(r[i], vo, so, v);
  }
  __device__ __forceinline__ void store_bits(int i, unsigned vo, unsigned so, unsigned short h) const {
    Stored<uint16_t>::store_bits(r[i], vo, so, h);
  }
};

// One resource for all nine planes, plane i at byte offset i * stride folded into the scalar
// offset: 4 SGPRs instead of 36 per buffer (higher occupancy: the plane resources otherwise
// make the step SGPR-limited). Valid when the nine planes span less than 4 GiB.
struct PlanesOne {
  Rsrc r;
  unsigned stride;
  template <class S>
  __device__ __forceinline__ void init(const S* base, size_t N) {
    r = make_rsrc(base);
    stride = static_cast<unsigned>(N * sizeof(S));
  }
  template <class S>
  __device__ __forceinline__ typename Stored<S>::C load(int i, unsigned vo, unsigned so) const {
    return Stored<S>::load(r, vo, so + static_cast<unsigned>(i) * stride);
  }
  template <class S>
  __device__ __forceinline__ typename WalkValue<S>::type load_v(int i, unsigned vo, unsigned so) const {
    if constexpr (std::is_same_v<S, uint16_t>)
      return Stored<uint16_t>::load_h(r, vo, so + static_cast<unsigned>(i) * stride);
    else
      return Stored<S>::load(r, vo, so + static_cast<unsigned>(i) * stride);
  }
  template <class S, int kAux>
  __device__ __forceinline__ void store(int i, unsigned vo, unsigned so, typename Stored<S>::C v) const {
    Stored<S>::template store<kAux>(r, vo, so + static_cast<unsigned>(i) * stride, v);
  }
  __device__ __forceinline__ void store_bits(int i, unsigned vo, unsigned so, unsigned short h) const {
    Stored<uint16_t>::store_bits(r, vo, so + static_cast<unsigned>(i) * stride, h);
  }
};

// Byte offsets of the three source rows (y - 1, y, y + 1) and columns (x - 1, x, x + 1).
struct Rows {
  unsigned m, c, p;
  __device__ __forceinline__ unsigned of(int cy) const { return cy > 0 ? m : (cy < 0 ? p : c); }
};
struct Cols {
  unsigned m, c, p;
  __device__ __forceinline__ unsigned of(int cx) const { return cx > 0 ? m : (cx < 0 ? p : c); }
};

// The nine pulled populations of a cell: direction i comes from (x - cx_i, y - cy_i).
template <class S, class PS = Planes, class C = typename Stored<S>::C>
__device__ __forceinline__ void pull_streamed(const PS& g, const Rows& rows, const Cols& cols,
                                              C (&v)[kQ]) {
#pragma unroll
  for (int i = 0; i < kQ; ++i) v[i] = g.template load<S>(i, cols.of(kCx[i]), rows.of(kCy[i]));
}

// The open-x inlet / outlet rules of one cell from its link code (solver_v3.cu:66-73,
// src/common/boundary.hpp:53-70), applied after its bounce-back corrections.
template <class S, class PS, class C>
__device__ __forceinline__ void apply_open_edges(uint32_t link, const PS& g, const Rows& rows,
                                                 const Cols& cols, const C* eq_in, C (&v)[kQ]) {
  if (link & kLinkInlet) {
    v[1] = eq_in[1];
    v[5] = eq_in[5];
    v[8] = eq_in[8];
  }
  if (link & kLinkOutlet) {
    // What x - 1 = nx - 2 receives in direction i: pulled from (nx - 1, y - cy_i), or
    // bounced back from opp(i) of cell nx - 2.
    const unsigned left = cols.c - static_cast<unsigned>(sizeof(S));
#pragma unroll
    for (int i : {3, 6, 7})
      v[i] = (link & outlet_bb_bit(i)) ? g.template load<S>(kOpp[i], left, rows.c)
                                       : g.template load<S>(i, cols.c, rows.of(kCy[i]));
  }
}

// Dirty-segment corrections for one cell, from its link code: bounce-back from the cell's own
// opposite slot (src/v3_cuda_opt/solver_v3.cu:74-75), then the open-x rules.
template <class S, bool kOpen, class PS = Planes, class C = typename Stored<S>::C>
__device__ __forceinline__ void apply_links(uint32_t link, const PS& g, const Rows& rows,
                                            const Cols& cols, const C* eq_in, C (&v)[kQ]) {
#pragma unroll
  for (int i = 1; i < kQ; ++i)
    if (link & (1u << i)) v[i] = g.template load<S>(kOpp[i], cols.c, rows.c);
  if constexpr (kOpen) apply_open_edges<S>(link, g, rows, cols, eq_in, v);
}

// Whole-domain gather (periodic y) of the post-streaming populations of fluid cell (x, y);
// returns false for a solid cell. Used by the read-back kernels.
template <class S, class C = typename Stored<S>::C>
__device__ __forceinline__ bool gather_cell(const S* g, const uint16_t* links, size_t N, int nx,
                                            int ny, const C* eq_in, int x, int y, C (&v)[kQ]) {
  const uint32_t link = links[static_cast<size_t>(y) * nx + x];
  if (link & kLinkSolid) return false;
  const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
  constexpr unsigned B = sizeof(S);
  const Rows rows{static_cast<unsigned>(ym * nx) * B, static_cast<unsigned>(y * nx) * B,
                  static_cast<unsigned>(yp * nx) * B};
  const Cols cols{static_cast<unsigned>(x == 0 ? nx - 1 : x - 1) * B, static_cast<unsigned>(x) * B,
                  static_cast<unsigned>(x == nx - 1 ? 0 : x + 1) * B};
  Planes planes;
  planes.init(g, N);
  pull_streamed<S>(planes, rows, cols, v);
  apply_links<S, true>(link, planes, rows, cols, eq_in, v);
  return true;
}

// Wavefront sum of a double with DPP row operations and the gfx9 row broadcasts; every lane of
// row 3 (lanes 48..63) ends with the total and lane 63's value is returned to all lanes. A fixed
// tree, so the result is deterministic.
template <int kCtrl, int kRowMask = 0xf, int kBankMask = 0xf, bool kBoundCtrl = false>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, kRowMask, kBankMask,
                                             kBoundCtrl);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, kRowMask, kBankMask,
                                             kBoundCtrl);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xb1>(v);        // quad_perm [1,0,3,2]
  v += dpp_f64<0x4e>(v);        // quad_perm [2,3,0,1]
  v += dpp_f64<0x141>(v);       // row_half_mirror: sums of 8
  v += dpp_f64<0x140>(v);       // row_mirror: sums of 16 in every lane of a row
  v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3: lane 63 holds the total
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                          __builtin_amdgcn_readlane(__double2loint(v), 63));
}

constexpr int kReduceThreads = 256;

// Block sums of 4 doubles (wave sums through LDS); thread 0 ends with them in s[0..3].
__device__ __forceinline__ void block_sum4(double (&s)[4]) {
  __shared__ double part[kReduceThreads / 64][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) s[k] = wave_sum(s[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k) part[wave][k] = s[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double t = 0.0;
      for (int w = 0; w < kReduceThreads / 64; ++w) t += part[w][k];
      s[k] = t;
    }
  }
}

// Second pass of the FP64 reductions (step.hip): sums the per-block partials in a fixed order.
__global__ void reduce_final_kernel(const double* __restrict__ partials, int nblocks,
                                    double* __restrict__ out);

}  // namespace lbmhip
