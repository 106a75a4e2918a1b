// D2Q9 lattice constants and the BGK + Guo collision, shared by host and device code.
//
// The collision performs exactly the floating-point operations of the reference's v0 cell
// update, in the same order (src/v0_cpu/solver_v0.cpp:26-39,72-83 with the helpers of
// src/common/d2q9.hpp:40-50): compiled with -ffp-contract=off and without fast-math, a float or
// double step is bit-identical to v0. The direction tables keep the reference numbering
// (src/common/d2q9.hpp:14-27) so that states and fixtures are interchangeable.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lbmhip {

constexpr int kQ = 9;
// 0 rest, 1 E, 2 N, 3 W, 4 S, 5 NE, 6 NW, 7 SW, 8 SE.
constexpr int kCx[kQ] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int kCy[kQ] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int kOpp[kQ] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

// Lattice weights, correctly rounded in T at compile time (4/9, 1/9, 1/36).
template <class T>
__host__ __device__ constexpr T weight(int i) {
  return i == 0 ? T(4) / T(9) : (i <= 4 ? T(1) / T(9) : T(1) / T(36));
}

// Per-cell link code (uint16), read only where a tile touches a wall or an open edge:
//   bit 0          the cell is solid (not updated);
//   bit i (1..8)   the pull source c - e_i is solid: direction i bounces back from opp(i) of c;
//   bit 9          open-x inlet column (directions 1, 5, 8 take the inlet equilibrium);
//   bit 10         open-x outlet column (directions 3, 6, 7 copy what x = nx - 2 receives);
//   bits 11,12,13  outlet bounce-back of directions 3, 6, 7 (source (nx-1, y - cy) is solid).
constexpr uint32_t kLinkSolid = 1u;
constexpr uint32_t kLinkInlet = 1u << 9;
constexpr uint32_t kLinkOutlet = 1u << 10;
__host__ __device__ constexpr uint32_t outlet_bb_bit(int i) {
  return i == 3 ? (1u << 11) : (i == 6 ? (1u << 12) : (1u << 13));
}

// Moments with the Guo half-force velocity, then BGK relaxation plus the Guo source, in place.
// Every product by a lattice constant is written as in the reference (int constant converted
// to T); the compiler may only apply IEEE-exact folds (1*x, -1*x, x + -y), never 0*x.
//
// Three rewrites of the reference's expressions cannot change a bit of any result that is
// finite (a non-finite state has no bitwise counterpart: NaN encodings differ between
// processors):
//  * the products 0 * g_i in the sums over c_x g_i and c_y g_i are left out: each running sum
//    starts at +0, and a round-to-nearest sum is -0 only when both operands are -0, so the sum
//    is never -0 and adding +-0 to it is the identity;
//  * 3 cu + 4.5 cu cu of the rest direction is left out: there cu = 0 * ux + 0 * uy = +-0, so
//    3 cu is +-0, 4.5 cu * cu is +0 and their sum is +0 whatever the signs; the bracket is
//    +0 - 1.5 usq;
//  * directions 3, 4, 7, 8 reuse the bracket terms of their opposites 1, 2, 5, 6: cu of a
//    direction and of its opposite are exact negatives unless both are zeros (round to
//    nearest is symmetric; the zero products only add +-0), so 4.5 cu * cu is the same value
//    and 3 cu the exact negative, making 3 cu + 4.5 cu cu = 4.5 cu cu - 3 cu of the opposite;
//    where cu is a zero both sums are +0, as above.
// Everything else, cu and cf in the source term included, keeps the reference's operations.
// All three are checked bit for bit against the reference's full expressions on random,
// signed-zero and subnormal inputs (tests/cpp/test_collide.cpp).
template <class T>
__host__ __device__ __forceinline__ void collide(T (&g)[kQ], T fx, T fy, T omega, T fp) {
  T m0 = T(0), mx = T(0), my = T(0);
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    m0 += g[i];
    if (kCx[i] != 0) mx += T(kCx[i]) * g[i];
    if (kCy[i] != 0) my += T(kCy[i]) * g[i];
  }
  const T drho = m0;
  const T rho = T(1) + drho;
  const T ux = (mx + T(0.5) * fx) / rho;
  const T uy = (my + T(0.5) * fy) / rho;
  const T usq = ux * ux + uy * uy;
  T cu[kQ], t3[kQ], t45[kQ];
#pragma unroll
  for (int i = 0; i < kQ; ++i) cu[i] = T(kCx[i]) * ux + T(kCy[i]) * uy;
#pragma unroll
  for (int i = 1; i < kQ; ++i) {
    if (i == 1 || i == 2 || i == 5 || i == 6) {
      t3[i] = T(3) * cu[i];
      t45[i] = T(4.5) * cu[i] * cu[i];
    }
  }
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    const T cxi = T(kCx[i]), cyi = T(kCy[i]), wi = weight<T>(i);
    const int o = kOpp[i];
    const bool first = i == 1 || i == 2 || i == 5 || i == 6;
    const T bracket = i == 0 ? T(0) - T(1.5) * usq
                      : first ? t3[i] + t45[i] - T(1.5) * usq
                              : t45[o] - t3[o] - T(1.5) * usq;
    const T geq = wi * (drho + rho * bracket);
    const T cf = cxi * fx + cyi * fy;
    const T src = wi * (T(3) * ((cxi - ux) * fx + (cyi - uy) * fy) + T(9) * cu[i] * cf);
    g[i] = g[i] - omega * (g[i] - geq) + fp * src;
  }
}

// Moments of nine post-streaming populations as the reference computes them for its fields
// (src/v3_cuda_opt/solver_v3.cu:322-343): rho = 1 + sum g, u = (sum c g + F/2) / rho.
template <class T>
__host__ __device__ __forceinline__ void moments(const T (&g)[kQ], T fx, T fy, T& rho, T& ux,
                                                 T& uy) {
  T m0 = T(0), mx = T(0), my = T(0);
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    m0 += g[i];
    mx += T(kCx[i]) * g[i];
    my += T(kCy[i]) * g[i];
  }
  rho = T(1) + m0;
  ux = (mx + T(0.5) * fx) / rho;
  uy = (my + T(0.5) * fy) / rho;
}

// Shifted equilibrium of a uniform state (rho0, ux0, uy0) (src/common/boundary.hpp:26-34).
template <class T>
__host__ __device__ inline void equilibrium_state(T rho0, T ux0, T uy0, T* eq) {
  const T usq = ux0 * ux0 + uy0 * uy0;
  const T drho0 = rho0 - T(1);
  for (int i = 0; i < kQ; ++i) {
    const T cu = T(kCx[i]) * ux0 + T(kCy[i]) * uy0;
    eq[i] = weight<T>(i) * (drho0 + rho0 * (T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq));
  }
}

}  // namespace lbmhip
