// Discrete adjoint of the fused step (no reference counterpart).
//
// Forward: R_{t+1}(c) = C(G(R_t)(c); F_c, tau) for fluid c, where G is the pull gather with
// halfway bounce-back (a 0/1 matrix: every fluid slot of R_t is read exactly once) and C the
// per-cell BGK + Guo collision with a per-cell force F_c. Objective:
//   J = k * sum_{t in window} sum_{c in probe} ux_c(t),  k = 1 / (|window| |probe|),
// with ux_c(t) = (sum_i cx_i gin_i + fx_c / 2) / rho computed from gin = G(R_t)(c).
//
// Reverse step t (one launch): for each fluid cell
//   lambda(c, i) = mu_{t+1}(c, opp i)        if c + e_i is solid (transposed bounce-back)
//                  mu_{t+1}(c + e_i, i)      otherwise (transpose of the pull = a push-side gather)
//   mu_t(c)      = DC(gin_t(c))^T lambda(c) + k dux/dgin  (direct term, probe cells in window)
//   dJ/dF_c     += dC/dF_c^T lambda(c) + k dux/dF_c;      dJ/dtau += dC/dtau^T lambda(c)
// The 9x9 Jacobian transpose is applied through the three moments (drho, ux, uy):
//   mu_j = (1 - omega) lambda_j + A_rho + cx_j A_mx + cy_j A_my,
// so a reverse step costs a gather of R_t, a gather of mu_{t+1}, and O(9) flops per direction.
// Periodic in x and y (walls and obstacles through the link codes).
#include "adjoint.hpp"
#include "step_body.hpp"

namespace lbmhip {
namespace {

static_assert(kAdjointThreads == kReduceThreads, "reverse blocks reduce with block_sum4");

// Reverse step for the 256 cells of reverse block b (one thread per cell).
template <class T>
__device__ __forceinline__ void adjoint_block(const AdjointParams<T>& p, unsigned b) {
  const size_t c = static_cast<size_t>(b) * kReduceThreads + threadIdx.x;
  T dtau = T(0);
  double jpart = 0.0;  // k * ux at probe cells of a state in the window: J, summed per block
  double acc0 = 0.0, acc1 = 0.0;  // this block's running sums, read early (no round trip at the end)
  if (threadIdx.x == 0) {
    acc0 = p.partials[2 * b];
    acc1 = p.partials[2 * b + 1];
  }
  if (c < p.N) {
    const int nx = p.nx, ny = p.ny;
    const int x = static_cast<int>(c % nx), y = static_cast<int>(c / nx);
    constexpr unsigned B = sizeof(T);
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    const Rows rows{static_cast<unsigned>(ym * nx) * B, static_cast<unsigned>(y * nx) * B,
                    static_cast<unsigned>(yp * nx) * B};
    const Cols cols{static_cast<unsigned>(x == 0 ? nx - 1 : x - 1) * B, static_cast<unsigned>(x) * B,
                    static_cast<unsigned>(x == nx - 1 ? 0 : x + 1) * B};
    // Every load of an interior cell is issued before the link code is examined (one memory
    // round trip per reverse step); bounce-back slots are patched afterwards, and a solid
    // cell's loads (all in range) are discarded.
    const uint32_t link = p.links[c];
    Planes R;
    R.init(p.R, p.N);
    T g[kQ];
    pull_streamed<T>(R, rows, cols, g);
    // lambda: mirror gather of mu_{t+1} (from c + e_i, i.e. the column/row of -c_i)
    T lam[kQ];
    Planes M;
    if (p.mu_in) {
      M.init(p.mu_in, p.N);
#pragma unroll
      for (int i = 0; i < kQ; ++i) lam[i] = Stored<T>::load(M.r[i], cols.of(-kCx[i]), rows.of(-kCy[i]));
    } else {
#pragma unroll
      for (int i = 0; i < kQ; ++i) lam[i] = T(0);
    }
    const T fx = p.force[c], fy = p.force[p.N + c];
    const T dfx0 = p.dforce[c], dfy0 = p.dforce[p.N + c];
    if (!(link & kLinkSolid)) {
#pragma unroll
      for (int i = 1; i < kQ; ++i)
        if (link & (1u << i)) g[i] = Stored<T>::load(R.r[kOpp[i]], cols.c, rows.c);
      if (p.mu_in)
#pragma unroll
        for (int i = 1; i < kQ; ++i)
          if (link & (1u << kOpp[i])) lam[i] = Stored<T>::load(M.r[kOpp[i]], cols.c, rows.c);

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
      const T om = p.omega, fp = p.fp;

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
      if (p.direct && x >= p.px0 && x < p.px0 + p.pw && y >= p.py0 && y < p.py0 + p.ph) {
        A_ux += p.k;
        jpart = p.kd * static_cast<double>(ux);
      }
      const T ir = T(1) / rho;
      const T A_mx = A_ux * ir, A_my = A_uy * ir;
      const T A_rho = A_drho - (A_ux * ux + A_uy * uy) * ir;
      Planes Mo;
      Mo.init(p.mu_out, p.N);
#pragma unroll
      for (int i = 0; i < kQ; ++i)
        Stored<T>::store(Mo.r[i], cols.c, rows.c,
                         (T(1) - om) * lam[i] + A_rho + T(kCx[i]) * A_mx + T(kCy[i]) * A_my);
      p.dforce[c] = dfx0 + (dfx + A_ux * T(0.5) * ir);
      p.dforce[p.N + c] = dfy0 + (dfy + A_uy * T(0.5) * ir);
      dtau = -(om * om) * dom;  // d omega / d tau = -omega^2
    }
  }
  // per-block partials of dJ/dtau and J, accumulated in FP64 into this block's own slots
  double s[4] = {static_cast<double>(dtau), jpart, 0.0, 0.0};
  block_sum4(s);
  if (threadIdx.x == 0) {
    p.partials[2 * b] = acc0 + s[0];
    p.partials[2 * b + 1] = acc1 + s[1];
  }
}

template <class T>
__global__ void __launch_bounds__(kReduceThreads) adjoint_step_kernel(const AdjointParams<T> p) {
  adjoint_block(p, blockIdx.x);
}

// One launch holding a reverse step and an independent forward step (different buffers): the
// first adjoint_blocks(N) workgroups run the reverse step, the rest one forward step of the
// recomputation chain (tiles of 256 columns x f.rows_per_tile rows, the same per-cell arithmetic
// as step_kernel, so the state is bit-identical). On small grids both launches are mostly
// latency; sharing one halves the launches of the checkpointed reverse sweep.
template <class T>
__global__ void __launch_bounds__(kReduceThreads) adjoint_fused_kernel(const AdjointParams<T> a,
                                                                       const StepParams<T> f) {
  const unsigned nrev = adjoint_blocks(a.N);
  if (blockIdx.x < nrev) {
    adjoint_block(a, blockIdx.x);
    return;
  }
  const unsigned tile = blockIdx.x - nrev;
  const int tx = static_cast<int>(tile % static_cast<unsigned>(f.tiles_x));
  const int ty = static_cast<int>(tile / static_cast<unsigned>(f.tiles_x));
  const int x = tx * kReduceThreads + static_cast<int>(threadIdx.x);
  const int y0 = ty * f.rows_per_tile;
  const int y1 = min(y0 + f.rows_per_tile, f.ny);
  const DomainView<T> d{f.g,  f.gn, f.links, f.segflags, f.force,  f.N,  f.nx,
                        f.ny, f.nseg, f.wrap_y, f.omega, f.fp,    f.fx, f.fy};
  update_column<T, false, true, kAuxNone, PlanesOne>(d, x, y0, y1, f.eq_in);
}

// J accumulation: k * sum over the probe of ux from the post-streaming state (one block).
template <class T>
__global__ void __launch_bounds__(256) probe_kernel(const GatherParams<T> p, int px0, int py0,
                                                    int pw, int ph, double k,
                                                    double* __restrict__ acc) {
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = threadIdx.x; j < pw * ph; j += blockDim.x) {
    const int x = px0 + j % pw, y = py0 + j / pw;
    T v[kQ];
    if (!gather_cell<T>(p.g, p.links, p.N, p.nx, p.ny, p.eq_in, x, y, v)) continue;
    const size_t c = static_cast<size_t>(y) * p.nx + x;
    T r, u, w;
    moments(v, p.force[c], p.force[p.N + c], r, u, w);
    s[0] += static_cast<double>(u);
  }
  block_sum4(s);
  if (threadIdx.x == 0) acc[0] += k * s[0];
}

template <class T>
__global__ void fill_kernel(T* __restrict__ a, size_t n, T v) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    a[i] = v;
}

}  // namespace

template <class T>
hipError_t launch_adjoint_step(const AdjointParams<T>& p, hipStream_t stream) {
  adjoint_step_kernel<T><<<adjoint_blocks(p.N), 256, 0, stream>>>(p);
  return hipGetLastError();
}
template hipError_t launch_adjoint_step<float>(const AdjointParams<float>&, hipStream_t);
template hipError_t launch_adjoint_step<double>(const AdjointParams<double>&, hipStream_t);

template <class T>
hipError_t launch_adjoint_fused(const AdjointParams<T>& a, const StepParams<T>& f,
                                hipStream_t stream) {
  // The forward part assumes 256-column tiles over the whole grid, periodic in y, one buffer
  // resource (nine planes < 4 GiB) and a per-cell force field.
  if (f.tiles_x != (f.nx + kReduceThreads - 1) / kReduceThreads || f.row0 != 0 ||
      f.nrows != f.ny || f.rows_per_tile < 1 || !f.wrap_y || !f.force || a.N != f.N ||
      static_cast<uint64_t>(kQ) * f.N * sizeof(T) >= (uint64_t{1} << 32))
    return hipErrorInvalidValue;
  const unsigned tiles_y = static_cast<unsigned>((f.ny + f.rows_per_tile - 1) / f.rows_per_tile);
  const unsigned grid = adjoint_blocks(a.N) + static_cast<unsigned>(f.tiles_x) * tiles_y;
  adjoint_fused_kernel<T><<<grid, kReduceThreads, 0, stream>>>(a, f);
  return hipGetLastError();
}
template hipError_t launch_adjoint_fused<float>(const AdjointParams<float>&,
                                                const StepParams<float>&, hipStream_t);
template hipError_t launch_adjoint_fused<double>(const AdjointParams<double>&,
                                                 const StepParams<double>&, hipStream_t);

template <class T>
hipError_t launch_probe(const GatherParams<T>& p, int px0, int py0, int pw, int ph, double k,
                        double* acc, hipStream_t stream) {
  probe_kernel<T><<<1, 256, 0, stream>>>(p, px0, py0, pw, ph, k, acc);
  return hipGetLastError();
}
template hipError_t launch_probe<float>(const GatherParams<float>&, int, int, int, int, double,
                                        double*, hipStream_t);
template hipError_t launch_probe<double>(const GatherParams<double>&, int, int, int, int, double,
                                         double*, hipStream_t);

template <class T>
hipError_t launch_fill(T* a, size_t n, T v, hipStream_t stream) {
  fill_kernel<T><<<256, 256, 0, stream>>>(a, n, v);
  return hipGetLastError();
}
template hipError_t launch_fill<float>(float*, size_t, float, hipStream_t);
template hipError_t launch_fill<double>(double*, size_t, double, hipStream_t);

}  // namespace lbmhip
