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

}  // namespace

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
