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

// The random word of a cell for one step. Steps 2p and 2p + 1 share the hash of (cell, seed of
// pair p): the even step reads it as is, the odd step rotated by 16 bits (shift = 16), so the
// rounding thresholds (bits 19-31 of each word, see to_half_sr) of the two steps come from
// disjoint bits of the hash. The two-step kernel hashes a cell once per launch.
__device__ __forceinline__ uint32_t step_random(uint32_t h, uint32_t shift) {
  return __builtin_amdgcn_alignbit(h, h, shift);
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
// The axis directions are rounded first; their momentum error (ex, ey) is taken off the
// diagonals by the mass-neutral shift -(ex c_x + ey c_y) / 4 before those are rounded; the rest
// population is then set from the cell's pre-collision mass (BGK and the Guo source conserve
// it) less the eight rounded moving values, so the stored mass changes by the rounding of that
// one value only. v[0] is not read. Rounding errors e = half(v) - v are exact in fp32 (v and its
// rounding are within a factor 2); each is one mixed-precision fma(half(v), one, -v) with
// one = 1 (see HalfCollide::unit), and so are the sums of rounded values.
__device__ __forceinline__ void round_cell(const float (&v)[kQ], float mass, uint32_t r,
                                           float one, unsigned short (&h)[kQ]) {
  uint32_t rnd[kQ];
#pragma unroll
  for (int i = 0; i < kQ; ++i) rnd[i] = r + static_cast<uint32_t>(i) * 0x9E3779B9u;
  float e[5];
#pragma unroll
  for (int i = 1; i <= 4; ++i) {
    h[i] = to_half_sr(v[i], rnd[i]);
    e[i] = __builtin_fmaf(half_value(h[i]), one, -v[i]);
  }
  const float ex = e[1] - e[3], ey = e[2] - e[4];
  // sum_d c_d (-(ex c_dx + ey c_dy) / 4) = -(ex, ey) over the diagonals d, and sum_d of the
  // shift = 0: directions 5 (1, 1), 6 (-1, 1), 7 (-1, -1), 8 (1, -1)
  const float sp = ex + ey, sm = ey - ex;
  h[5] = to_half_sr(__builtin_fmaf(-0.25f, sp, v[5]), rnd[5]);
  h[6] = to_half_sr(__builtin_fmaf(-0.25f, sm, v[6]), rnd[6]);
  h[7] = to_half_sr(__builtin_fmaf(0.25f, sp, v[7]), rnd[7]);
  h[8] = to_half_sr(__builtin_fmaf(0.25f, sm, v[8]), rnd[8]);
  // mass - sum_{i >= 1} half(h_i), in independent partial sums
  const float a = __builtin_fmaf(half_value(h[1]), -one, mass);
  const float b = __builtin_fmaf(half_value(h[2]), one, half_value(h[3]));
  const float c = __builtin_fmaf(half_value(h[4]), one, half_value(h[5]));
  const float d = __builtin_fmaf(half_value(h[6]), one, half_value(h[7]));
  const float rest = __builtin_fmaf(half_value(h[8]), -one, a - ((b + c) + d));
  h[0] = to_half_sr(rest, rnd[0]);
}

// Weight class of direction i: 0 rest, 1 axis, 2 diagonal.
__device__ __forceinline__ constexpr int wclass(int i) { return i == 0 ? 0 : (i < 5 ? 1 : 2); }

// BGK + Guo in scaled units with the constants of HalfCollide: per pair of opposite directions
// one even and one odd part (7 operations for two directions). With b = omega w S rho the even
// part b (4.5 cu^2 - 1.5 u.u) + B_i cu + omega w S drho - C u.F is evaluated as
// (4.5 b cu + B_i) cu + e_k, its direction-independent terms e_k gathered once per weight class:
// fp32 rounding differs from forming the polynomial first by far less than the binary16
// storage resolves. Every fused multiply-add is written out (no contraction left to the
// compiler), so the one-step and two-step FP16 kernels perform the same operations on a cell;
// sums with an operand that may come straight from binary16 storage are fma(a, unit, b).
// Returns the cell's mass S drho (the sum of its nine pulled values). v[0] is left as it was:
// round_cell sets the rest population from the mass.
__device__ __forceinline__ float fast_collide(const HalfCollide& c, float (&v)[kQ]) {
  const float one = c.unit;
  auto add = [one](float a, float b) { return __builtin_fmaf(a, one, b); };
  auto sub = [one](float a, float b) { return __builtin_fmaf(a, one, -b); };
  const float m0 = add(v[8], (add(v[0], v[1]) + add(v[2], v[3])) + (add(v[4], v[5]) + add(v[6], v[7])));
  const float d57 = sub(v[5], v[7]), d86 = sub(v[8], v[6]);
  const float mx = (sub(v[1], v[3]) + d57) + d86;
  const float my = (sub(v[2], v[4]) + d57) - d86;
  const float rho = __builtin_fmaf(m0, c.invS, 1.0f);
  // u = (m / S + F / 2) / rho = (m + S F / 2) (1 / (S rho)): S a power of two, so the same
  // value as fma(m, 1 / S, F / 2) / rho, with one scalar operand per instruction
  const float irS = __builtin_amdgcn_rcpf(rho) * c.invS;
  const float ux = (mx + c.hfxS) * irS;
  const float uy = (my + c.hfyS) * irS;
  const float rusq = rho * __builtin_fmaf(ux, ux, uy * uy);
  const float uF = __builtin_fmaf(ux, c.fx, uy * c.fy);
  // per moving weight class (0 axis, 1 diagonal; the rest direction is not updated here):
  // e_k = omega w S (drho - 1.5 rho u.u) - C u.F, and 4.5 and 3 omega w S rho
  float e[2], b45[2], b3[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    e[k] = __builtin_fmaf(-c.owS15[k + 1], rusq, __builtin_fmaf(c.ow[k + 1], m0, -(c.C[k + 1] * uF)));
    b45[k] = c.owS45[k + 1] * rho;
    b3[k] = c.owS3[k + 1] * rho;
  }
  // pairs (direction, opposite) with the direction's cu: (1, 3) ux, (2, 4) uy, (5, 7) ux + uy,
  // (6, 8) uy - ux
  const float spp = ux + uy, smp = uy - ux;
  constexpr int kFirst[4] = {1, 2, 5, 6};
  const float cus[4] = {ux, uy, spp, smp};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int i = kFirst[j], o = kOpp[i], k = wclass(i) - 1;
    const float cu = cus[j];
    const float even = __builtin_fmaf(__builtin_fmaf(b45[k], cu, c.B[i]), cu, e[k]);
    const float odd = __builtin_fmaf(b3[k], cu, c.A[i]);
    v[i] = __builtin_fmaf(c.om1, v[i], even + odd);
    v[o] = __builtin_fmaf(c.om1, v[o], even - odd);
  }
  return m0;
}

}  // namespace lbmhip
