// Momentum-exchange force on solid bodies inside a box (drag and lift for the cylinder
// instance; the reference samples a velocity probe instead, src/common/driver.hpp:186-251).
//
// With halfway bounce-back, the population leaving fluid cell c towards a solid neighbour
// c + e_i is the stored post-collision value R_i(c); it returns as opp(i) the next step, so the
// body receives 2 (R_i(c) + w_i) c_i per link and step (f = g + w). Summed in FP64 over the fluid
// cells of the box in a fixed order (one block, DPP wave sums), so the signal is deterministic.
#include "device_common.hpp"
#include "kernels.hpp"

namespace lbmhip {
namespace {

template <class S>
__global__ void __launch_bounds__(kReduceThreads)
    exchange_kernel(const S* __restrict__ g, const uint16_t* __restrict__ links, size_t N, int nx,
                    int x0, int y0, int w, int h, float inv_scale, double* __restrict__ out) {
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  Planes planes;
  planes.init(g, N);
  for (int j = threadIdx.x; j < w * h; j += kReduceThreads) {
    const int x = x0 + j % w, y = y0 + j / w;
    const size_t c = static_cast<size_t>(y) * nx + x;
    const uint32_t link = links[c];
    if (link & kLinkSolid) continue;
    const unsigned vo = static_cast<unsigned>(x * sizeof(S)), so = static_cast<unsigned>(y * nx * sizeof(S));
#pragma unroll
    for (int i = 1; i < kQ; ++i) {
      if (!(link & (1u << kOpp[i]))) continue;  // c + e_i is not solid
      double f = static_cast<double>(Stored<S>::load(planes.r[i], vo, so));
      if constexpr (sizeof(S) == 2) f *= inv_scale;
      f = 2.0 * (f + weight<double>(i));
      s[0] += kCx[i] * f;
      s[1] += kCy[i] * f;
    }
  }
  block_sum4(s);
  if (threadIdx.x == 0) {
    out[0] = s[0];
    out[1] = s[1];
  }
}

// Vorticity of a velocity field (src/common/io.cpp:54-74): central differences, periodic in y;
// in x periodic, or one-sided (step 1) at x = 0 and x = nx - 1 of an open channel; 0 in solid
// cells. The same operations in T as the reference's host loop.
template <class T>
__global__ void vorticity_kernel(const T* __restrict__ ux, const T* __restrict__ uy,
                                 const uint16_t* __restrict__ links, int nx, int ny,
                                 int periodic_x, T* __restrict__ w) {
  const size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= static_cast<size_t>(nx) * ny) return;
  if (links[c] & kLinkSolid) {
    w[c] = T(0);
    return;
  }
  const int x = static_cast<int>(c % nx), y = static_cast<int>(c / nx);
  const int yp = y == ny - 1 ? 0 : y + 1, ym = y == 0 ? ny - 1 : y - 1;
  int xp = x == nx - 1 ? 0 : x + 1, xm = x == 0 ? nx - 1 : x - 1;
  T h = T(2);
  if (!periodic_x && x == 0) {
    xm = 0;
    h = T(1);
  }
  if (!periodic_x && x == nx - 1) {
    xp = nx - 1;
    h = T(1);
  }
  const size_t row = static_cast<size_t>(y) * nx;
  const T duy_dx = (uy[row + xp] - uy[row + xm]) / h;
  const T dux_dy = (ux[static_cast<size_t>(yp) * nx + x] - ux[static_cast<size_t>(ym) * nx + x]) / T(2);
  w[c] = duy_dx - dux_dy;
}

}  // namespace

template <class T>
hipError_t launch_vorticity(const T* ux, const T* uy, const uint16_t* links, int nx, int ny,
                            int periodic_x, T* w, hipStream_t stream) {
  const size_t cells = static_cast<size_t>(nx) * ny;
  if (cells == 0) return hipSuccess;
  vorticity_kernel<T><<<static_cast<unsigned>((cells + 255) / 256), 256, 0, stream>>>(
      ux, uy, links, nx, ny, periodic_x, w);
  return hipGetLastError();
}
template hipError_t launch_vorticity<float>(const float*, const float*, const uint16_t*, int, int,
                                            int, float*, hipStream_t);
template hipError_t launch_vorticity<double>(const double*, const double*, const uint16_t*, int,
                                             int, int, double*, hipStream_t);

template <class S>
hipError_t launch_exchange(const S* g, const uint16_t* links, size_t N, int nx, int x0, int y0,
                           int w, int h, float inv_scale, double* out, hipStream_t stream) {
  exchange_kernel<S><<<1, kReduceThreads, 0, stream>>>(g, links, N, nx, x0, y0, w, h, inv_scale, out);
  return hipGetLastError();
}
template hipError_t launch_exchange<float>(const float*, const uint16_t*, size_t, int, int, int,
                                           int, int, float, double*, hipStream_t);
template hipError_t launch_exchange<double>(const double*, const uint16_t*, size_t, int, int, int,
                                            int, int, float, double*, hipStream_t);
template hipError_t launch_exchange<uint16_t>(const uint16_t*, const uint16_t*, size_t, int, int,
                                              int, int, int, float, double*, hipStream_t);

}  // namespace lbmhip
