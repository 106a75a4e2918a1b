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

// Forward step of one cell: the same values and the same collide() as step_kernel (the state is
// bit-identical); optionally accumulates k * ux of the pulled state for J.
template <class T>
__device__ __forceinline__ void forward_cell(const AdjGeom<T>& G, const T* src, T* dst,
                                             double* jacc, unsigned c, int x, int y, bool dirty) {
  const unsigned stride = static_cast<unsigned>(G.N * sizeof(T));
  const Nbr nb = neighbours<T>(x, y, G.nx, G.ny);
  const Rsrc rs = make_rsrc(src);
  T v[kQ], own[kQ];
#pragma unroll
  for (int i = 0; i < kQ; ++i) v[i] = Stored<T>::load(rs, nb.pull(i), i * stride);
  uint32_t link = 0;
  if (dirty) {
    link = G.links[c];
#pragma unroll
    for (int i = 1; i < kQ; ++i) own[i] = Stored<T>::load(rs, nb.self(), i * stride);
  }
  const T fx = G.force[c], fy = G.force[G.N + c];
  const bool probe = jacc && x >= G.px0 && x < G.px0 + G.pw && y >= G.py0 && y < G.py0 + G.ph;
  const unsigned pj = static_cast<unsigned>((y - G.py0) * G.pw + (x - G.px0));
  double j0 = 0.0;
  if (probe) j0 = jacc[pj];
  if (link & kLinkSolid) return;
  if (dirty)
#pragma unroll
    for (int i = 1; i < kQ; ++i)
      if (link & (1u << i)) v[i] = own[kOpp[i]];
  if (probe) {
    T r, u, w;
    moments(v, fx, fy, r, u, w);
    jacc[pj] = j0 + G.kd * static_cast<double>(u);
  }
  if (!dst) return;
  collide(v, fx, fy, G.omega, G.fp);
  const Rsrc rd = make_rsrc(dst);
#pragma unroll
  for (int i = 0; i < kQ; ++i) Stored<T>::store(rd, nb.self(), i * stride, v[i]);
}

// Reverse step of one cell.
template <class T>
__device__ __forceinline__ void reverse_cell(const AdjGeom<T>& G, const AdjRev<T>& p, unsigned c,
                                             int x, int y, bool dirty) {
  const unsigned stride = static_cast<unsigned>(G.N * sizeof(T));
  const Nbr nb = neighbours<T>(x, y, G.nx, G.ny);
  const Rsrc rr = make_rsrc(p.R);
  const Rsrc rm = make_rsrc(p.mu_in ? p.mu_in : p.R);  // never read when mu_in is null
  const bool have_mu = p.mu_in != nullptr;
  // every load of the cell is issued before any is used: one memory round trip
  T g[kQ], lam[kQ], ownR[kQ], ownM[kQ];
#pragma unroll
  for (int i = 0; i < kQ; ++i) g[i] = Stored<T>::load(rr, nb.pull(i), i * stride);
  if (have_mu) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) lam[i] = Stored<T>::load(rm, nb.push(i), i * stride);
  } else {
#pragma unroll
    for (int i = 0; i < kQ; ++i) lam[i] = T(0);
  }
  uint32_t link = 0;
  if (dirty) {
    link = G.links[c];
#pragma unroll
    for (int i = 1; i < kQ; ++i) ownR[i] = Stored<T>::load(rr, nb.self(), i * stride);
    if (have_mu)
#pragma unroll
      for (int i = 1; i < kQ; ++i) ownM[i] = Stored<T>::load(rm, nb.self(), i * stride);
  }
  const T fx = G.force[c], fy = G.force[G.N + c];
  const T dfx0 = p.dforce[c], dfy0 = p.dforce[G.N + c];
  const double t0 = p.tacc[c];
  const bool probe =
      p.direct && x >= G.px0 && x < G.px0 + G.pw && y >= G.py0 && y < G.py0 + G.ph;
  const unsigned pj = static_cast<unsigned>((y - G.py0) * G.pw + (x - G.px0));
  double j0 = 0.0;
  if (probe && p.jacc) j0 = p.jacc[pj];
  if (link & kLinkSolid) return;
  if (dirty) {
#pragma unroll
    for (int i = 1; i < kQ; ++i)
      if (link & (1u << i)) g[i] = ownR[kOpp[i]];
    if (have_mu)
#pragma unroll
      for (int i = 1; i < kQ; ++i)
        if (link & (1u << kOpp[i])) lam[i] = ownM[kOpp[i]];
  }

  T drho = T(0), mx = T(0), my = T(0);
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    drho += g[i];
    mx += T(kCx[i]) * g[i];
    my += T(kCy[i]) * g[i];
  }
  const T rho = T(1) + drho;
  const T ux = (mx + T(0.5) * fx) / rho, uy = (my + T(0.5) * fy) / rho;
  const T usq = ux * ux + uy * uy;
  const T om = G.omega, fp = G.fp;

  T A_drho = T(0), A_ux = T(0), A_uy = T(0), dfx = T(0), dfy = T(0), dom = T(0);
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    const T cx = T(kCx[i]), cy = T(kCy[i]), w = weight<T>(i);
    const T cu = cx * ux + cy * uy;
    const T P = T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq;
    const T geq = w * (drho + rho * P);
    const T cf = cx * fx + cy * fy;
    const T src = w * (T(3) * ((cx - ux) * fx + (cy - uy) * fy) + T(9) * cu * cf);
    const T l = lam[i];
    A_drho += l * (om * w * (T(1) + P));
    A_ux += l * (om * w * rho * (T(3) * cx + T(9) * cu * cx - T(3) * ux) +
                 fp * w * (T(-3) * fx + T(9) * cx * cf));
    A_uy += l * (om * w * rho * (T(3) * cy + T(9) * cu * cy - T(3) * uy) +
                 fp * w * (T(-3) * fy + T(9) * cy * cf));
    dfx += l * (fp * w * (T(3) * (cx - ux) + T(9) * cu * cx));
    dfy += l * (fp * w * (T(3) * (cy - uy) + T(9) * cu * cy));
    dom += l * (geq - g[i] - T(0.5) * src);
  }
  // direct term of J: k * ux at probe cells, states inside the averaging window
  if (probe) {
    A_ux += G.k;
    if (p.jacc) p.jacc[pj] = j0 + G.kd * static_cast<double>(ux);
  }
  const T ir = T(1) / rho;
  const T A_mx = A_ux * ir, A_my = A_uy * ir;
  const T A_rho = A_drho - (A_ux * ux + A_uy * uy) * ir;
  const Rsrc ro = make_rsrc(p.mu_out);
#pragma unroll
  for (int i = 0; i < kQ; ++i)
    Stored<T>::store(ro, nb.self(), i * stride,
                     (T(1) - om) * lam[i] + A_rho + T(kCx[i]) * A_mx + T(kCy[i]) * A_my);
  p.dforce[c] = dfx0 + (dfx + A_ux * T(0.5) * ir);
  p.dforce[G.N + c] = dfy0 + (dfy + A_uy * T(0.5) * ir);
  const T dtau = -(om * om) * dom;  // d omega / d tau = -omega^2
  p.tacc[c] = t0 + static_cast<double>(dtau);
}

// Blocks [0, nb) run the reverse step (when has_rev), the next nb blocks each forward step.
template <class T>
__global__ void __launch_bounds__(kAdjointThreads)
    adjoint_kernel(const AdjGeom<T> G, const AdjRev<T> r, const AdjFwd<T> f0, const AdjFwd<T> f1,
                   const AdjFwd<T> f2, int has_rev) {
  const unsigned nb = adjoint_blocks(G.N);
  unsigned role = blockIdx.x / nb;
  const unsigned c = (blockIdx.x - role * nb) * kAdjointThreads + threadIdx.x;
  if (c >= G.N) return;
  const unsigned nx = static_cast<unsigned>(G.nx);
  const int y = static_cast<int>(c / nx), x = static_cast<int>(c - static_cast<unsigned>(y) * nx);
  const bool dirty = in_boxes(G.boxes, x, y);
  if (has_rev) {
    if (role == 0) {
      reverse_cell(G, r, c, x, y, dirty);
      return;
    }
    --role;
  }
  const T* src = role == 0 ? f0.src : (role == 1 ? f1.src : f2.src);
  T* dst = role == 0 ? f0.dst : (role == 1 ? f1.dst : f2.dst);
  double* jacc = role == 0 ? f0.jacc : (role == 1 ? f1.jacc : f2.jacc);
  forward_cell(G, src, dst, jacc, c, x, y, dirty);
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
  AdjFwd<T> f[kAdjMaxFwd] = {};
  for (int k = 0; k < nfwd; ++k) f[k] = fwd[k];
  const unsigned grid = adjoint_blocks(g.N) * static_cast<unsigned>((has_rev ? 1 : 0) + nfwd);
  adjoint_kernel<T><<<grid, kAdjointThreads, 0, stream>>>(g, rev, f[0], f[1], f[2], has_rev ? 1 : 0);
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
