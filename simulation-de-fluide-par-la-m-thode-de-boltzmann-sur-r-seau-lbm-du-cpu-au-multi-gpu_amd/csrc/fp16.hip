// FP16-storage form of the D2Q9 step: populations stored as IEEE binary16 of g * S (S a power of
// two), all arithmetic in fp32. No reference counterpart: it halves the bytes per cell update
// (36 B instead of 72 B) of the same stencil as step.hip.
//
// Arithmetic (not bit-compatible with v0; checked against the FP64 step):
//   * Moments and collision in scaled units (g' = S g): S drho = sum g', and the equilibrium
//     w (S drho + rho S (3 cu + 4.5 cu^2 - 1.5 u^2)) needs no per-population rescaling.
//   * Explicit FMAs, one reciprocal of rho; omega w and omega w S rho folded per weight class,
//     so each direction costs the polynomial, one FMA for omega geq, and the Guo source
//     fma(B_i, c_i.u, A_i - C_k u.F); the constants come precomputed with the launch
//     (half_cell.hpp, shared with the two-step kernel).
//   * Stores round stochastically to binary16 with the hardware conversion v_cvt_sr_f16_f32
//     (its rounding threshold comes from the high bits of a 32-bit random operand): a hash of
//     the cell and the step pair (steps 2p and 2p + 1 read it as is and rotated by 16 bits),
//     advanced by the golden-ratio increment for each of the nine values.
//     Round-to-nearest would be biased the same way in every cell of a uniform region and would
//     stall the slow diffusive modes once their per-step change drops below half an ulp; the
//     unbiased rounding keeps the expected update exact.
//   * The nine roundings of a cell are chained so that its mass and momentum carry only the
//     noise of the finest values: the axis directions are rounded first and their momentum
//     error is taken off the diagonals (a mass-neutral shift), then the rest population is set
//     to the cell's pre-collision mass less the eight rounded moving values (BGK and the Guo
//     source conserve the mass), so the stored mass changes only by that value's rounding.
//     Conserved moments integrate rounding
//     noise without damping (total mass, and the x-staggered momentum mode u ~ (-1)^(x+t),
//     which D2Q9 streaming and BGK leave invariant for any tau), so this cuts the mass drift
//     by two orders of magnitude and the staggered-mode noise about threefold.
#include "device_common.hpp"
#include "half_cell.hpp"
#include "kernels.hpp"

namespace lbmhip {
namespace {

template <bool kOpen, bool kField, bool kOne>
__global__ void __launch_bounds__(1024) step_fp16_kernel(const StepParamsH p) {
  using PS = std::conditional_t<kOne, PlanesOne, Planes>;
  const unsigned tile = tile_in_order(p.xcd_bands, blockIdx.x, gridDim.x, static_cast<unsigned>(p.tiles_x));
  const int tx = static_cast<int>(tile % static_cast<unsigned>(p.tiles_x));
  const int ty = static_cast<int>(tile / static_cast<unsigned>(p.tiles_x));
  const int nx = p.nx;
  const int x = tx * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
  if (__builtin_amdgcn_readfirstlane(x) >= nx) return;  // whole wave past the last column
  const bool in_x = x < nx;
  const unsigned xc = in_x ? x : nx - 1;
  constexpr unsigned B = sizeof(uint16_t);
  const Cols cols{(xc == 0 ? nx - 1 : xc - 1) * B, xc * B,
                  (xc + 1 == static_cast<unsigned>(nx) ? 0 : xc + 1) * B};
  const int seg = __builtin_amdgcn_readfirstlane(x) >> 6;
  const int y0 = p.row0 + ty * p.rows_per_tile;
  const int y1 = min(y0 + p.rows_per_tile, p.row0 + p.nrows);
  PS g, gn;
  g.init(p.g, p.N);
  gn.init(p.gn, p.N);

  for (int y = y0; y < y1; ++y) {
    int ym = y - 1, yp = y + 1;
    if (p.wrap_y) {
      ym = ym < 0 ? p.ny - 1 : ym;
      yp = yp >= p.ny ? 0 : yp;
    }
    const Rows rows{static_cast<unsigned>(ym * nx) * B, static_cast<unsigned>(y * nx) * B,
                    static_cast<unsigned>(yp * nx) * B};
    float v[kQ];
    pull_streamed<uint16_t>(g, rows, cols, v);
    bool active = in_x;
    if (p.segflags[static_cast<size_t>(y) * p.nseg + seg]) {
      const uint32_t link = in_x ? p.links[static_cast<size_t>(y) * nx + xc] : kLinkSolid;
      active = !(link & kLinkSolid);
      if (active) apply_links<uint16_t, kOpen>(link, g, rows, cols, p.eq_in, v);
    }
    if (active) {
      const uint32_t cell = static_cast<uint32_t>(y) * static_cast<uint32_t>(nx) + xc;
      float mass;
      if constexpr (kField) {
        HalfCollide c2;
        c2.setup(p.omega, p.fp, p.scale, p.inv_scale, p.force[cell], p.force[p.N + cell]);
        mass = fast_collide(c2, v);
      } else {
        mass = fast_collide(p.col, v);
      }
      unsigned short h[kQ];
      round_cell(v, mass, step_random(hash_cell(cell, p.seed), p.shift), p.col.unit, h);
#pragma unroll
      for (int i = 0; i < kQ; ++i) gn.store_bits(i, cols.c, rows.c, h[i]);
    }
  }
}

// ---- read-back (values returned unscaled, in float) ------------------------------------------

__device__ __forceinline__ bool gather_h(const GatherParamsH& p, int x, int y,
                                         float (&v)[kQ]) {
  if (!gather_cell<uint16_t>(p.g, p.links, p.N, p.nx, p.ny, p.eq_in_s, x, y, v)) return false;
#pragma unroll
  for (int i = 0; i < kQ; ++i) v[i] *= p.inv_scale;
  return true;
}

__global__ void gather_fp16_kernel(const GatherParamsH p, float* __restrict__ out) {
  const size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= static_cast<size_t>(p.nx) * p.ny) return;
  float v[kQ];
  if (!gather_h(p, static_cast<int>(c % p.nx), static_cast<int>(c / p.nx), v))
#pragma unroll
    for (int i = 0; i < kQ; ++i) v[i] = 0.0f;
#pragma unroll
  for (int i = 0; i < kQ; ++i) out[c * kQ + i] = v[i];
}

__global__ void gather_one_fp16_kernel(const GatherParamsH p, int x, int y,
                                       float* __restrict__ out) {
  float v[kQ];
  if (!gather_h(p, x, y, v))
#pragma unroll
    for (int i = 0; i < kQ; ++i) v[i] = 0.0f;
#pragma unroll
  for (int i = 0; i < kQ; ++i) out[i] = v[i];
}

__global__ void macro_fp16_kernel(const GatherParamsH p, float fx, float fy,
                                  const float* __restrict__ force, float* __restrict__ rho,
                                  float* __restrict__ ux, float* __restrict__ uy) {
  const size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= static_cast<size_t>(p.nx) * p.ny) return;
  float v[kQ];
  float r = 0.0f, u = 0.0f, w = 0.0f;
  if (gather_h(p, static_cast<int>(c % p.nx), static_cast<int>(c / p.nx), v)) {
    if (force) {
      fx = force[c];
      fy = force[p.N + c];
    }
    moments(v, fx, fy, r, u, w);
  }
  rho[c] = r;
  ux[c] = u;
  uy[c] = w;
}

__global__ void __launch_bounds__(kReduceThreads)
    reduce_fp16_kernel(const GatherParamsH p, double* __restrict__ partials) {
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  const size_t cells = static_cast<size_t>(p.nx) * p.ny;
  for (size_t c = static_cast<size_t>(blockIdx.x) * kReduceThreads + threadIdx.x; c < cells;
       c += static_cast<size_t>(gridDim.x) * kReduceThreads) {
    float v[kQ];
    if (!gather_h(p, static_cast<int>(c % p.nx), static_cast<int>(c / p.nx), v)) continue;
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

}  // namespace

hipError_t launch_step_fp16(const StepParamsH& p, bool open_x, int block_x, hipStream_t stream) {
  const int tiles_y = (p.nrows + p.rows_per_tile - 1) / p.rows_per_tile;
  const unsigned grid = static_cast<unsigned>(p.tiles_x) * static_cast<unsigned>(tiles_y);
  if (grid == 0) return hipSuccess;
  const bool field = p.force != nullptr;
  // One buffer resource (4 SGPRs instead of 36 per buffer) when the nine planes span < 4 GiB.
  const bool one = static_cast<uint64_t>(kQ) * p.N * sizeof(uint16_t) < (uint64_t{1} << 32);
  if (open_x) {
    if (field)
      step_fp16_kernel<true, true, false><<<grid, block_x, 0, stream>>>(p);
    else
      step_fp16_kernel<true, false, false><<<grid, block_x, 0, stream>>>(p);
  } else if (field) {
    step_fp16_kernel<false, true, false><<<grid, block_x, 0, stream>>>(p);
  } else if (one) {
    step_fp16_kernel<false, false, true><<<grid, block_x, 0, stream>>>(p);
  } else {
    step_fp16_kernel<false, false, false><<<grid, block_x, 0, stream>>>(p);
  }
  return hipGetLastError();
}

hipError_t launch_gather_fp16(const GatherParamsH& p, float* out, hipStream_t stream) {
  const size_t cells = static_cast<size_t>(p.nx) * p.ny;
  gather_fp16_kernel<<<static_cast<unsigned>((cells + 255) / 256), 256, 0, stream>>>(p, out);
  return hipGetLastError();
}

hipError_t launch_gather_one_fp16(const GatherParamsH& p, int x, int y, float* out9,
                                  hipStream_t stream) {
  gather_one_fp16_kernel<<<1, 1, 0, stream>>>(p, x, y, out9);
  return hipGetLastError();
}

hipError_t launch_macro_fp16(const GatherParamsH& p, float fx, float fy,
                             const float* force, float* rho, float* ux, float* uy,
                             hipStream_t stream) {
  const size_t cells = static_cast<size_t>(p.nx) * p.ny;
  macro_fp16_kernel<<<static_cast<unsigned>((cells + 255) / 256), 256, 0, stream>>>(p, fx, fy, force,
                                                                                  rho, ux, uy);
  return hipGetLastError();
}

hipError_t launch_reduce_fp16(const GatherParamsH& p, double* partials, double* out,
                              hipStream_t stream) {
  reduce_fp16_kernel<<<kReduceBlocks, kReduceThreads, 0, stream>>>(p, partials);
  reduce_final_kernel<<<1, kReduceThreads, 0, stream>>>(partials, kReduceBlocks, out);
  return hipGetLastError();
}

}  // namespace lbmhip
