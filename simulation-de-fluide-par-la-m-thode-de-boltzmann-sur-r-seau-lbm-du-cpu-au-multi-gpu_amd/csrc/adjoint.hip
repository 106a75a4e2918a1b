// Discrete adjoint of the fused step (no reference counterpart).
//
// Forward: R_{t+1}(c) = C(G(R_t)(c); F_c, tau) for fluid c, where G is the pull gather with
// halfway bounce-back (a 0/1 matrix: every fluid slot of R_t is read exactly once) and C the
// per-cell BGK + Guo collision with a per-cell force F_c. Objective:
//   J = k * sum_{t in window} sum_{c in probe} ux_c(t),  k = 1 / (|window| |probe|),
// with ux_c(t) = (sum_i cx_i gin_i + fx_c / 2) / rho computed from gin = G(R_t)(c).
//
// Reverse step t: for each fluid cell
//   lambda(c, i) = mu_{t+1}(c, opp i)        if c + e_i is solid (transposed bounce-back)
//                  mu_{t+1}(c + e_i, i)      otherwise (transpose of the pull = a push-side gather)
//   mu_t(c)      = DC(gin_t(c))^T lambda(c) + k dux/dgin  (direct term, probe cells in window)
//   dJ/dF_c     += dC/dF_c^T lambda(c) + k dux/dF_c;      dJ/dtau += dC/dtau^T lambda(c)
// The 9x9 Jacobian transpose is applied through the three moments (drho, ux, uy):
//   mu_j = (1 - omega) lambda_j + A_rho + cx_j A_mx + cy_j A_my,
// so a reverse step costs a gather of R_t, a gather of mu_{t+1}, and O(9) flops per direction.
// dJ/dtau and J accumulate per cell in FP64 (no block reduction, no barrier: every step is one
// memory round trip); one fixed-order sum at the end makes them deterministic.
// Periodic in x and y (walls and obstacles through the link codes).
#include "adjoint.hpp"
#include "step_body.hpp"

namespace lbmhip {
namespace {

__device__ __forceinline__ bool in_boxes(const DirtyBoxes& b, int x, int y) {
  bool in = false;
#pragma unroll
  for (int k = 0; k < kMaxBoxes; ++k)
    in |= (k < b.n) & (x >= b.x0[k]) & (x < b.x1[k]) & (y >= b.y0[k]) & (y < b.y1[k]);
  return in;
}

// Byte offsets (within a plane) of the cell and of its eight neighbours (periodic).
struct Nbr {
  Rows rows;
  Cols cols;
  __device__ __forceinline__ unsigned pull(int i) const {  // source of direction i: c - e_i
    return cols.of(kCx[i]) + rows.of(kCy[i]);
  }
  __device__ __forceinline__ unsigned push(int i) const {  // c + e_i
    return cols.of(-kCx[i]) + rows.of(-kCy[i]);
  }
  __device__ __forceinline__ unsigned self() const { return cols.c + rows.c; }
};

template <class T>
__device__ __forceinline__ Nbr neighbours(int x, int y, int nx, int ny) {
  constexpr unsigned B = sizeof(T);
  const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
  return Nbr{Rows{static_cast<unsigned>(ym * nx) * B, static_cast<unsigned>(y * nx) * B,
                  static_cast<unsigned>(yp * nx) * B},
             Cols{static_cast<unsigned>(x == 0 ? nx - 1 : x - 1) * B, static_cast<unsigned>(x) * B,
                  static_cast<unsigned>(x == nx - 1 ? 0 : x + 1) * B}};
}

// Makes the compiler hold the value of x at this point (an empty asm with x as an in/out
// register operand): loads issued before it cannot be moved below it. The kernels pin every
// loaded value after their load phase, so all loads of a launch are in flight before the first
// wait, instead of the population loads being sunk below the solid-cell test.
template <class T>
__device__ __forceinline__ void pin(T& x) {
  asm volatile("" : "+v"(x));
}
template <class T, int n>
__device__ __forceinline__ void pin(T (&x)[n], int from = 0) {
#pragma unroll
  for (int i = from; i < n; ++i) pin(x[i]);
}

// Values of one cell shared by every role of a launch (the same cell, mask and force field).
template <class T>
struct CellIn {
  uint32_t link;
  T fx, fy;
};

// Forward step of one cell, split into a load phase and a compute/store phase so that a launch
// can issue the loads of all its roles before it waits for any of them. The values and the
// collide() are those of step_kernel: the state is bit-identical.
template <class T>
struct FwdIn {
  T v[kQ], own[kQ];
  double j0;
  __device__ __forceinline__ void pin_all() {
    pin(v);
    pin(own, 1);
    pin(j0);
  }
};

template <class T>
__device__ __forceinline__ void forward_load(const T* src, double* jacc, const Nbr& nb,
                                             unsigned stride, bool dirty, bool probe, unsigned pj,
                                             FwdIn<T>& in) {
  const Rsrc rs = make_rsrc(src);
#pragma unroll
  for (int i = 0; i < kQ; ++i) in.v[i] = Stored<T>::load(rs, nb.pull(i), i * stride);
#pragma unroll
  for (int i = 1; i < kQ; ++i) in.own[i] = T(0);
  if (dirty)  // wave-uniform
#pragma unroll
    for (int i = 1; i < kQ; ++i) in.own[i] = Stored<T>::load(rs, nb.self(), i * stride);
  in.j0 = 0.0;
  if (jacc) in.j0 = jacc[probe ? pj : 0u];  // uniform branch, in-range index for every lane
}

template <class T>
__device__ __forceinline__ void forward_finish(const AdjGeom<T>& G, T* dst, double* jacc,
                                               const Nbr& nb, unsigned stride, bool dirty,
                                               bool probe, unsigned pj, const CellIn<T>& cell,
                                               FwdIn<T>& in) {
#pragma unroll
  for (int i = 1; i < kQ; ++i)
    if (cell.link & (1u << i)) in.v[i] = in.own[kOpp[i]];
  if (probe && jacc) {
    T r, u, w;
    moments(in.v, cell.fx, cell.fy, r, u, w);
    jacc[pj] = in.j0 + G.kd * static_cast<double>(u);
  }
  if (!dst) return;
  collide(in.v, cell.fx, cell.fy, G.omega, G.fp);
  const Rsrc rd = make_rsrc(dst);
#pragma unroll
  for (int i = 0; i < kQ; ++i) Stored<T>::store(rd, nb.self(), i * stride, in.v[i]);
}

// Reverse step of one cell, in the same two phases.
template <class T>
struct RevIn {
  T g[kQ], lam[kQ], ownR[kQ], ownM[kQ];
  T dfx0, dfy0;
  double t0, j0;
  __device__ __forceinline__ void pin_all() {
    pin(g);
    pin(lam);
    pin(ownR, 1);
    pin(ownM, 1);
    pin(dfx0);
    pin(dfy0);
    pin(t0);
    pin(j0);
  }
};

template <class T>
__device__ __forceinline__ void reverse_load(const AdjGeom<T>& G, const AdjRev<T>& p, unsigned c,
                                             const Nbr& nb, unsigned stride, bool dirty,
                                             bool probe, unsigned pj, RevIn<T>& in) {
  const Rsrc rr = make_rsrc(p.R);
  const bool have_mu = p.mu_in != nullptr;
  const Rsrc rm = make_rsrc(have_mu ? p.mu_in : p.R);  // never read when mu_in is null
#pragma unroll
  for (int i = 0; i < kQ; ++i) in.g[i] = Stored<T>::load(rr, nb.pull(i), i * stride);
  if (have_mu) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) in.lam[i] = Stored<T>::load(rm, nb.push(i), i * stride);
  } else {
#pragma unroll
    for (int i = 0; i < kQ; ++i) in.lam[i] = T(0);
  }
#pragma unroll
  for (int i = 1; i < kQ; ++i) in.ownR[i] = in.ownM[i] = T(0);
  if (dirty) {  // wave-uniform
#pragma unroll
    for (int i = 1; i < kQ; ++i) in.ownR[i] = Stored<T>::load(rr, nb.self(), i * stride);
    if (have_mu)
#pragma unroll
      for (int i = 1; i < kQ; ++i) in.ownM[i] = Stored<T>::load(rm, nb.self(), i * stride);
  }
  in.dfx0 = p.dforce[c];
  in.dfy0 = p.dforce[G.N + c];
  in.t0 = p.tacc[c];
  in.j0 = 0.0;
  if (p.jacc) in.j0 = p.jacc[probe ? pj : 0u];  // uniform branch, in-range index for every lane
}

template <class T>
__device__ __forceinline__ void reverse_finish(const AdjGeom<T>& G, const AdjRev<T>& p, unsigned c,
                                               const Nbr& nb, unsigned stride, bool dirty,
                                               bool probe, unsigned pj, const CellIn<T>& cell,
                                               RevIn<T>& in) {
#pragma clang fp contract(fast)  // the reverse step has no bitwise counterpart
  const T fx = cell.fx, fy = cell.fy;
#pragma unroll
  for (int i = 1; i < kQ; ++i)
    if (cell.link & (1u << i)) in.g[i] = in.ownR[kOpp[i]];
  if (p.mu_in)
#pragma unroll
    for (int i = 1; i < kQ; ++i)
      if (cell.link & (1u << kOpp[i])) in.lam[i] = in.ownM[kOpp[i]];
  const T(&g)[kQ] = in.g;
  const T(&lam)[kQ] = in.lam;
  // The transposed Jacobian in moment form.
  // With a_i = lambda_i w_i and its moments S0 = sum a, Sx = sum cx a, Sy = sum cy a,
  // Sxx = sum cx^2 a, Syy = sum cy^2 a, Sxy = sum cx cy a (cx, cy in {-1, 0, 1}), every
  // direction sum of the BGK + Guo Jacobian is a polynomial in them:
  //   sum a P        = 3 (ux Sx + uy Sy) + 4.5 (ux^2 Sxx + 2 ux uy Sxy + uy^2 Syy) - 1.5 |u|^2 S0
  //   sum a cu cx    = ux Sxx + uy Sxy        sum a cu cy = ux Sxy + uy Syy
  //   sum a cx cf    = fx Sxx + fy Sxy        sum a cy cf = fx Sxy + fy Syy
  //   sum a cu cf    = ux fx Sxx + (ux fy + uy fx) Sxy + uy fy Syy
  const T drho = ((g[0] + g[1]) + (g[2] + g[3])) + ((g[4] + g[5]) + (g[6] + g[7])) + g[8];
  const T mx = ((g[1] - g[3]) + (g[5] - g[6])) + (g[8] - g[7]);
  const T my = ((g[2] - g[4]) + (g[5] + g[6])) - (g[7] + g[8]);
  const T rho = T(1) + drho;
  const T ir = T(1) / rho;
  const T ux = (mx + T(0.5) * fx) * ir, uy = (my + T(0.5) * fy) * ir;
  const T usq = ux * ux + uy * uy;
  const T om = G.omega, fp = G.fp;
  const T w0 = weight<T>(0), w1 = weight<T>(1), w5 = weight<T>(5);
  const T a0 = w0 * lam[0];
  const T a1 = w1 * lam[1], a2 = w1 * lam[2], a3 = w1 * lam[3], a4 = w1 * lam[4];
  const T a5 = w5 * lam[5], a6 = w5 * lam[6], a7 = w5 * lam[7], a8 = w5 * lam[8];
  const T dsum = (a5 + a6) + (a7 + a8);  // diagonals
  const T S0 = a0 + ((a1 + a2) + (a3 + a4)) + dsum;
  const T Sx = (a1 - a3) + ((a5 - a6) - (a7 - a8));
  const T Sy = (a2 - a4) + ((a5 + a6) - (a7 + a8));
  const T Sxx = (a1 + a3) + dsum;
  const T Syy = (a2 + a4) + dsum;
  const T Sxy = (a5 - a6) + (a7 - a8);
  T Slg = T(0);
#pragma unroll
  for (int i = 0; i < kQ; ++i) Slg += lam[i] * g[i];
  const T Acu = ux * Sx + uy * Sy;
  const T AP = T(3) * Acu + T(4.5) * (ux * ux * Sxx + T(2) * ux * uy * Sxy + uy * uy * Syy) -
               T(1.5) * usq * S0;
  const T Ax = ux * Sxx + uy * Sxy, Ay = ux * Sxy + uy * Syy;
  const T Fx = fx * Sxx + fy * Sxy, Fy = fx * Sxy + fy * Syy;
  const T A_drho = om * (S0 + AP);
  T A_ux = om * rho * (T(3) * Sx + T(9) * Ax - T(3) * ux * S0) + fp * (T(-3) * fx * S0 + T(9) * Fx);
  const T A_uy = om * rho * (T(3) * Sy + T(9) * Ay - T(3) * uy * S0) + fp * (T(-3) * fy * S0 + T(9) * Fy);
  const T dfx = fp * (T(3) * (Sx - ux * S0) + T(9) * Ax);
  const T dfy = fp * (T(3) * (Sy - uy * S0) + T(9) * Ay);
  const T Acf = ux * fx * Sxx + (ux * fy + uy * fx) * Sxy + uy * fy * Syy;
  const T Lsrc = T(3) * ((fx * Sx + fy * Sy) - (ux * fx + uy * fy) * S0) + T(9) * Acf;
  const T dom = (drho * S0 + rho * AP) - Slg - T(0.5) * Lsrc;
  // direct term of J: k * ux at probe cells, states inside the averaging window; J itself
  // accumulates the forward's own ux (moments() in the step's operation order, no contraction),
  // so the gradient sweep reports the same J as the forward run
  if (probe) {
    A_ux += G.k;
    if (p.jacc) {
      T r, u, w;
      moments(g, fx, fy, r, u, w);
      p.jacc[pj] = in.j0 + G.kd * static_cast<double>(u);
    }
  }
  const T A_mx = A_ux * ir, A_my = A_uy * ir;
  const T A_rho = A_drho - (A_ux * ux + A_uy * uy) * ir;
  const Rsrc ro = make_rsrc(p.mu_out);
#pragma unroll
  for (int i = 0; i < kQ; ++i)
    Stored<T>::store(ro, nb.self(), i * stride,
                     (T(1) - om) * lam[i] + A_rho + T(kCx[i]) * A_mx + T(kCy[i]) * A_my);
  p.dforce[c] = in.dfx0 + (dfx + A_ux * T(0.5) * ir);
  p.dforce[G.N + c] = in.dfy0 + (dfy + A_uy * T(0.5) * ir);
  const T dtau = -(om * om) * dom;  // d omega / d tau = -omega^2
  p.tacc[c] = in.t0 + static_cast<double>(dtau);
}

// Tile of block b among the nb blocks of one role: each of the 8 XCDs walks a contiguous band
// of tiles (neighbouring rows meet in its L2) when nb is a multiple of 8, so that block b of
// every role runs on XCD b % 8.
__device__ __forceinline__ unsigned tile_of(unsigned b, unsigned nb) {
  return nb % kXcds == 0 ? xcd_tile(b, nb) : b;
}

// Where a thread of tile t sits, and what every role needs of its cell. A tile is
// kAdjointThreads columns of one row, so the row index is uniform over the block.
template <class T>
struct CellAt {
  int x, y;
  unsigned c, pj, stride;
  bool dirty, probe;
  Nbr nb;
  CellIn<T> cell;
};

// Returns false for a lane past the last column of the row.
template <class T>
__device__ __forceinline__ bool cell_at(const AdjGeom<T>& G, unsigned t, CellAt<T>& a) {
  const unsigned tiles_x = (static_cast<unsigned>(G.nx) + kAdjointThreads - 1) / kAdjointThreads;
  const unsigned ty = t / tiles_x;
  a.y = static_cast<int>(ty);
  a.x = static_cast<int>((t - ty * tiles_x) * kAdjointThreads + threadIdx.x);
  if (a.x >= G.nx) return false;
  a.c = static_cast<unsigned>(a.y) * static_cast<unsigned>(G.nx) + static_cast<unsigned>(a.x);
  // Waves with a cell in a dirty box load the link codes and the cells' own populations (for
  // every lane: valid addresses, and a clean cell's link code is 0); the others skip them.
  a.dirty = __builtin_amdgcn_ballot_w64(in_boxes(G.boxes, a.x, a.y)) != 0;
  a.probe = (a.x >= G.px0) & (a.x < G.px0 + G.pw) & (a.y >= G.py0) & (a.y < G.py0 + G.ph);
  a.pj = static_cast<unsigned>((a.y - G.py0) * G.pw + (a.x - G.px0));
  a.stride = static_cast<unsigned>(G.N * sizeof(T));
  a.nb = neighbours<T>(a.x, a.y, G.nx, G.ny);
  a.cell.link = 0u;
  if (a.dirty) a.cell.link = G.links[a.c];
  a.cell.fx = G.force[a.c];
  a.cell.fy = G.force[G.N + a.c];
  return true;
}

template <class T>
__device__ __forceinline__ void forward_role(const AdjGeom<T>& G, CellAt<T>& a, const T* src,
                                             T* dst, double* jacc) {
  FwdIn<T> in;
  forward_load(src, jacc, a.nb, a.stride, a.dirty, a.probe, a.pj, in);
  pin(a.cell.link);
  pin(a.cell.fx);
  pin(a.cell.fy);
  in.pin_all();
  if (a.cell.link & kLinkSolid) return;
  forward_finish(G, dst, jacc, a.nb, a.stride, a.dirty, a.probe, a.pj, a.cell, in);
}

template <class T>
__device__ __forceinline__ void reverse_role(const AdjGeom<T>& G, CellAt<T>& a, const AdjRev<T>& r) {
  const bool pr = a.probe && r.direct;
  RevIn<T> in;
  reverse_load(G, r, a.c, a.nb, a.stride, a.dirty, pr, a.pj, in);
  pin(a.cell.link);
  pin(a.cell.fx);
  pin(a.cell.fy);
  in.pin_all();
  if (a.cell.link & kLinkSolid) return;
  reverse_finish(G, r, a.c, a.nb, a.stride, a.dirty, pr, a.pj, a.cell, in);
}

// Blocks [0, nb) run the reverse step (kRev), the next nb blocks each forward step: the roles
// run in different waves, each with a single memory round trip.
template <class T, int kRev>
__global__ void __launch_bounds__(kAdjointThreads) adjoint_kernel(const AdjArgs<T> a) {
  const AdjGeom<T>& G = a.g;
  const unsigned nb = adjoint_blocks(G.nx, G.ny);
  const unsigned role = blockIdx.x / nb;
  CellAt<T> at;
  if (!cell_at(G, tile_of(blockIdx.x - role * nb, nb), at)) return;
  if (kRev && role == 0) return reverse_role(G, at, a.rev);
  const unsigned k = role - (kRev ? 1u : 0u);
  const T* src = k == 0 ? a.fwd[0].src : (k == 1 ? a.fwd[1].src : a.fwd[2].src);
  T* dst = k == 0 ? a.fwd[0].dst : (k == 1 ? a.fwd[1].dst : a.fwd[2].dst);
  double* jacc = k == 0 ? a.fwd[0].jacc : (k == 1 ? a.fwd[1].jacc : a.fwd[2].jacc);
  forward_role(G, at, src, dst, jacc);
}

// A launch holding one forward step only (the forward run, the first pass of the gradient):
// the same per-cell code with a smaller argument block.
template <class T>
__global__ void __launch_bounds__(kAdjointThreads) adjoint_forward_kernel(const AdjGeom<T> G,
                                                                          const AdjFwd<T> f) {
  CellAt<T> at;
  if (!cell_at(G, tile_of(blockIdx.x, adjoint_blocks(G.nx, G.ny)), at)) return;
  forward_role(G, at, f.src, f.dst, f.jacc);
}

__global__ void __launch_bounds__(kReduceThreads) sum_f64_kernel(const double* __restrict__ a,
                                                                 size_t n, double* __restrict__ out) {
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (size_t i = threadIdx.x; i < n; i += kReduceThreads) s[0] += a[i];
  block_sum4(s);
  if (threadIdx.x == 0) out[0] = s[0];
}

template <class T>
__global__ void fill_kernel(T* __restrict__ a, size_t n, T v) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    a[i] = v;
}

}  // namespace

template <class T>
hipError_t launch_adjoint(const AdjGeom<T>& g, const AdjRev<T>& rev, bool has_rev,
                          const AdjFwd<T>* fwd, int nfwd, hipStream_t stream) {
  // one buffer resource per state (nine planes < 4 GiB), 32-bit cell indices
  if (nfwd < 0 || nfwd > kAdjMaxFwd || (!has_rev && nfwd == 0) || g.N == 0 ||
      static_cast<uint64_t>(kQ) * g.N * sizeof(T) >= (uint64_t{1} << 32) || g.boxes.n < 0 ||
      g.boxes.n > kMaxBoxes)
    return hipErrorInvalidValue;
  AdjArgs<T> a{};
  a.g = g;
  a.rev = rev;
  for (int k = 0; k < nfwd; ++k) a.fwd[k] = fwd[k];
  const unsigned grid = adjoint_blocks(g.nx, g.ny) * static_cast<unsigned>((has_rev ? 1 : 0) + nfwd);
  if (has_rev)
    adjoint_kernel<T, 1><<<grid, kAdjointThreads, 0, stream>>>(a);
  else if (nfwd == 1)
    adjoint_forward_kernel<T><<<grid, kAdjointThreads, 0, stream>>>(g, a.fwd[0]);
  else
    adjoint_kernel<T, 0><<<grid, kAdjointThreads, 0, stream>>>(a);
  return hipGetLastError();
}
template hipError_t launch_adjoint<float>(const AdjGeom<float>&, const AdjRev<float>&, bool,
                                          const AdjFwd<float>*, int, hipStream_t);
template hipError_t launch_adjoint<double>(const AdjGeom<double>&, const AdjRev<double>&, bool,
                                           const AdjFwd<double>*, int, hipStream_t);

hipError_t launch_sum_f64(const double* a, size_t n, double* out, hipStream_t stream) {
  sum_f64_kernel<<<1, kReduceThreads, 0, stream>>>(a, n, out);
  return hipGetLastError();
}

template <class T>
hipError_t launch_fill(T* a, size_t n, T v, hipStream_t stream) {
  fill_kernel<T><<<256, 256, 0, stream>>>(a, n, v);
  return hipGetLastError();
}
template hipError_t launch_fill<float>(float*, size_t, float, hipStream_t);
template hipError_t launch_fill<double>(double*, size_t, double, hipStream_t);

}  // namespace lbmhip
