// Cell routines of the FP16-storage form, shared by the one-step kernel (fp16.hip) and the
// two-step kernel (step2.hip): the collision in scaled units and the moment-compensated
// stochastic rounding of a cell's nine post-collision values (see fp16.hip for the method).
#pragma once

#include "device_common.hpp"
#include "kernels.hpp"

namespace lbmhip {

__device__ __forceinline__ uint32_t hash_cell(uint32_t c, uint32_t seed) {
  uint32_t h = (c ^ seed) * 0x9E3779B1u;
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

// fp32 (scaled units) -> binary16 bits, stochastic rounding driven by the high bits of r
// (bits 19-31 decide, measured). The instruction writes the low half of its destination and
// leaves the high half as it was. Through the builtin the compiler clears a register for that
// kept half before every conversion (9 extra instructions per cell update); written as an
// ALU instruction with a plain output, the high half is simply unspecified, and nothing reads
// it: only the low 16 bits are returned, and their users (16-bit stores, f16 -> f32
// conversions) read only those. Same bits as the builtin; the two-step FP16 kernel runs 3.8 %
// faster. (Writing a second value into the kept half through the builtin's op_sel form was
// tried: this compiler's placement of that half disagreed with the hardware's.)
__device__ __forceinline__ unsigned short to_half_sr(float v, uint32_t r) {
  uint32_t out;
  asm("v_cvt_sr_f16_f32 %0, %1, %2" : "=v"(out) : "v"(v), "v"(r));
  return static_cast<unsigned short>(out);
}

__device__ __forceinline__ float half_value(unsigned short h) {
  return static_cast<float>(__builtin_bit_cast(_Float16, h));
}

// Moment-compensated stochastic rounding of one cell's post-collision values (scaled units).
// Rounding errors e = half(v) - v are exact in fp32 (v and its rounding are within a factor 2).
__device__ __forceinline__ void round_cell(const float (&v)[kQ], uint32_t r,
                                           unsigned short (&h)[kQ]) {
  uint32_t rnd[kQ];
#pragma unroll
  for (int i = 0; i < kQ; ++i) rnd[i] = r + static_cast<uint32_t>(i) * 0x9E3779B9u;
  float e[kQ];
#pragma unroll
  for (int i = 1; i <= 4; ++i) {
    h[i] = to_half_sr(v[i], rnd[i]);
    e[i] = half_value(h[i]) - v[i];
  }
  const float ex = e[1] - e[3], ey = e[2] - e[4];
  float em = (e[1] + e[3]) + (e[2] + e[4]);
#pragma unroll
  for (int i = 5; i < kQ; ++i) {  // sum_d c_d (-(ex c_dx + ey c_dy) / 4) = -(ex, ey), sum_d = 0
    const float t = v[i] - 0.25f * (kCx[i] * ex + kCy[i] * ey);
    h[i] = to_half_sr(t, rnd[i]);
    em += half_value(h[i]) - v[i];
  }
  h[0] = to_half_sr(v[0] - em, rnd[0]);
}

// Weight class of direction i: 0 rest, 1 axis, 2 diagonal.
__device__ __forceinline__ constexpr int wclass(int i) { return i == 0 ? 0 : (i < 5 ? 1 : 2); }

// BGK + Guo in scaled units with the constants of HalfCollide. Every fused multiply-add is
// written out (no contraction left to the compiler), so the one-step and two-step FP16 kernels
// perform the same operations on a cell.
__device__ __forceinline__ void fast_collide(const HalfCollide& c, float (&v)[kQ]) {
  const float m0 = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7])) + v[8];
  const float d57 = v[5] - v[7], d86 = v[8] - v[6];
  const float mx = ((v[1] - v[3]) + d57) + d86;
  const float my = ((v[2] - v[4]) + d57) - d86;
  const float rho = __builtin_fmaf(m0, c.invS, 1.0f);
  const float ir = __builtin_amdgcn_rcpf(rho);
  const float ux = __builtin_fmaf(mx, c.invS, c.hfx) * ir;
  const float uy = __builtin_fmaf(my, c.invS, c.hfy) * ir;
  const float usq15 = 1.5f * __builtin_fmaf(ux, ux, uy * uy);
  const float uF = __builtin_fmaf(ux, c.fx, uy * c.fy);
  float a[3], b[3], cuf[3];  // per class: omega w m0, omega w S rho, C_k u.F
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    a[k] = c.ow[k] * m0;
    b[k] = c.owS[k] * rho;
    cuf[k] = c.C[k] * uF;
  }
  // c_i . u per direction (0, ux, uy, -ux, -uy, ux + uy, uy - ux, -(ux + uy), ux - uy)
  const float spp = ux + uy, smp = uy - ux;
  const float cus[kQ] = {0.0f, ux, uy, -ux, -uy, spp, smp, -spp, -smp};
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    const int k = wclass(i);
    const float cu = cus[i];
    const float poly = __builtin_fmaf(__builtin_fmaf(4.5f, cu, 3.0f), cu, -usq15);
    const float ogeq = __builtin_fmaf(b[k], poly, a[k]);
    const float src = __builtin_fmaf(c.B[i], cu, c.A[i] - cuf[k]);
    v[i] = __builtin_fmaf(c.om1, v[i], ogeq) + src;
  }
}

}  // namespace lbmhip
