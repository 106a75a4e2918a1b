// Two fused time steps per launch (temporal blocking), periodic domain with a uniform force:
// R_{t+2} from R_t with every population read once from HBM and written once, i.e. 72 B (fp32)
// per two cell updates instead of per one. Each cell update performs exactly the operations of
// step_kernel (the same pull, the same bounce-back rule, the same collide()), so the result is
// bit-identical to two single steps (src/v3_cuda_opt/solver_v3.cu:45-113 applied twice).
//
// A workgroup owns a tile of W = blockDim.x columns by H rows of the output and walks its rows
// downwards (in the +y direction). At row r it
//   1. computes step 1 (R_t -> R_{t+1}) at row r for its W columns, pulling from HBM (rows r - 1,
//      r, r + 1 of R_t; the row loads of neighbouring iterations hit L2);
//   2. publishes the six populations its neighbour columns pull in step 2 (planes with c_x != 0)
//      in a four-row ring in LDS;
//   3. computes step 2 (R_{t+1} -> R_{t+2}) at row r - 1: planes with c_x = 0 come from the
//      thread's own registers (rows r - 2, r - 1, r), the others from the ring (columns x - 1,
//      x + 1), bounce-back from the thread's own R_{t+1} values of row r - 1; and stores it.
// Step 1 therefore runs on rows y0 - 1 .. y1 (one extra row on each side of the tile) and the
// two columns just outside the tile (x0 - 1 and x0 + W) are computed once per row in a prologue
// (only the three populations the tile pulls from each of them are kept, in LDS). One barrier
// per row: with four ring slots a slot is rewritten two barriers after its last read.
#include "kernels.hpp"
#include "step_body.hpp"

namespace lbmhip {
namespace {

constexpr int kRingSlots = 4;
// Planes published in the ring and their slot index: the ones neighbours pull (c_x != 0).
__host__ __device__ constexpr int ring_index(int plane) {
  return plane == 1 ? 0 : plane == 3 ? 1 : plane == 5 ? 2 : plane == 6 ? 3 : plane == 7 ? 4 : 5;
}
constexpr int kRingPlanes = 6;
// Planes kept of the halo columns: the left one (x0 - 1) gives the tile its c_x = +1 planes,
// the right one (x0 + W) its c_x = -1 planes.
constexpr int kLeftPlanes[3] = {1, 5, 8};
constexpr int kRightPlanes[3] = {3, 6, 7};

// Step 1 at one cell (c, r) (both wrapped): the pull with its boundary corrections and the
// collision, exactly as update_column. Returns the link code (bit 0: solid, then v is
// meaningless).
template <class T>
__device__ __forceinline__ uint32_t first_step(const StepParams<T>& p, const PlanesOne& g, int c,
                                               int r, T (&v)[kQ]) {
  constexpr unsigned B = sizeof(T);
  const int nx = p.nx, ny = p.ny;
  const int rm = r == 0 ? ny - 1 : r - 1, rp = r == ny - 1 ? 0 : r + 1;
  const Rows rows{static_cast<unsigned>(rm * nx) * B, static_cast<unsigned>(r * nx) * B,
                  static_cast<unsigned>(rp * nx) * B};
  const Cols cols{static_cast<unsigned>(c == 0 ? nx - 1 : c - 1) * B, static_cast<unsigned>(c) * B,
                  static_cast<unsigned>(c == nx - 1 ? 0 : c + 1) * B};
  pull_streamed<T>(g, rows, cols, v);
  uint32_t link = 0;
  if (p.segflags[static_cast<size_t>(r) * p.nseg + (c >> 6)]) {
    link = p.links[static_cast<size_t>(r) * nx + c];
    if (!(link & kLinkSolid)) apply_links<T, false>(link, g, rows, cols, static_cast<const T*>(nullptr), v);
  }
  if (!(link & kLinkSolid)) collide(v, p.fx, p.fy, p.omega, p.fp);
  return link;
}

template <class T, int kStoreAux>
__global__ void __launch_bounds__(512) step2_kernel(const StepParams<T> p) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int W = static_cast<int>(blockDim.x);
  const int H = p.rows_per_tile;
  T* ring = reinterpret_cast<T*>(smem);                 // [kRingSlots][kRingPlanes][W]
  T* halo = ring + kRingSlots * kRingPlanes * W;        // [2][H + 2][3]
  const int t = static_cast<int>(threadIdx.x);
  const int nx = p.nx, ny = p.ny;
  const int tx = static_cast<int>(blockIdx.x % static_cast<unsigned>(p.tiles_x));
  const int ty = static_cast<int>(blockIdx.x / static_cast<unsigned>(p.tiles_x));
  const int x0 = tx * W, wv = min(W, nx - x0);  // valid columns of this tile
  const int y0 = ty * H, y1 = min(y0 + H, ny);
  const int nr = y1 - y0 + 2;                   // step-1 rows y0 - 1 .. y1
  PlanesOne g, gn;
  g.init(p.g, p.N);
  gn.init(p.gn, p.N);
  auto wrap_y = [&](int r) { return r < 0 ? r + ny : (r >= ny ? r - ny : r); };

  // prologue: step 1 on the two halo columns for every step-1 row of the tile
  for (int j = t; j < 2 * nr; j += W) {
    const int side = j >= nr ? 1 : 0, k = j - side * nr;
    const int c = side ? (x0 + wv == nx ? 0 : x0 + wv) : (x0 == 0 ? nx - 1 : x0 - 1);
    T v[kQ];
    first_step(p, g, c, wrap_y(y0 - 1 + k), v);
    T* h = halo + (side * (H + 2) + k) * 3;
#pragma unroll
    for (int q = 0; q < 3; ++q) h[q] = v[side ? kRightPlanes[q] : kLeftPlanes[q]];
  }
  __syncthreads();

  const bool active = t < wv;
  const int c = active ? x0 + t : x0 + wv - 1;  // idle lanes shadow the last column
  constexpr unsigned B = sizeof(T);
  T prev[kQ];           // R_{t+1} at (c, r - 1)
  T prev2 = T(0);       // plane 2 of R_{t+1} at (c, r - 2)
  uint32_t prev_link = kLinkSolid;
#pragma unroll
  for (int i = 0; i < kQ; ++i) prev[i] = T(0);
  for (int k = 0; k < nr; ++k) {
    const int r = wrap_y(y0 - 1 + k);
    T v[kQ];
    const uint32_t link = first_step(p, g, c, r, v);
    T* slot = ring + (k & (kRingSlots - 1)) * kRingPlanes * W;
#pragma unroll
    for (int i : {1, 3, 5, 6, 7, 8}) slot[ring_index(i) * W + t] = v[i];
    __syncthreads();
    if (k >= 2 && active && !(prev_link & kLinkSolid)) {
      // step 2 at row q = r - 1 (an output row of the tile)
      const T* s0 = ring + (k & (kRingSlots - 1)) * kRingPlanes * W;        // row q + 1
      const T* s1 = ring + ((k - 1) & (kRingSlots - 1)) * kRingPlanes * W;  // row q
      const T* s2 = ring + ((k - 2) & (kRingSlots - 1)) * kRingPlanes * W;  // row q - 1
      const T* hl = halo + k * 3;                // left halo, row q + 1 (k), q (k - 1), q - 1
      const T* hr = halo + ((H + 2) + k) * 3;
      const bool first = t == 0, last = t == wv - 1;
      const int tl = first ? 0 : t - 1, tr = last ? t : t + 1;  // in-range LDS columns
      T w[kQ];
      w[0] = prev[0];
      w[1] = first ? hl[-3 + 0] : s1[ring_index(1) * W + tl];
      w[2] = prev2;
      w[3] = last ? hr[-3 + 0] : s1[ring_index(3) * W + tr];
      w[4] = v[4];
      w[5] = first ? hl[-6 + 1] : s2[ring_index(5) * W + tl];
      w[6] = last ? hr[-6 + 1] : s2[ring_index(6) * W + tr];
      w[7] = last ? hr[2] : s0[ring_index(7) * W + tr];
      w[8] = first ? hl[2] : s0[ring_index(8) * W + tl];
#pragma unroll
      for (int i = 1; i < kQ; ++i)
        if (prev_link & (1u << i)) w[i] = prev[kOpp[i]];
      collide(w, p.fx, p.fy, p.omega, p.fp);
      const int q = y0 - 2 + k;
      const unsigned vo = static_cast<unsigned>(c) * B, so = static_cast<unsigned>(q * nx) * B;
#pragma unroll
      for (int i = 0; i < kQ; ++i) gn.template store<T, kStoreAux>(i, vo, so, w[i]);
    }
    prev2 = prev[2];
#pragma unroll
    for (int i = 0; i < kQ; ++i) prev[i] = v[i];
    prev_link = link;
  }
}

}  // namespace

size_t step2_lds_bytes(int block_x, int rows, size_t elem) {
  return (static_cast<size_t>(kRingSlots) * kRingPlanes * block_x + 2 * static_cast<size_t>(rows + 2) * 3) * elem;
}

template <class T>
hipError_t launch_step2(const StepParams<T>& p, int block_x, hipStream_t stream) {
  // periodic, uniform force, one buffer resource, whole domain, tile rows >= 1
  if (!p.wrap_y || p.force || p.row0 != 0 || p.nrows != p.ny || p.rows_per_tile < 1 ||
      block_x < 64 || block_x > 512 || block_x % 64 ||
      p.tiles_x != (p.nx + block_x - 1) / block_x ||
      static_cast<uint64_t>(kQ) * p.N * sizeof(T) >= (uint64_t{1} << 32))
    return hipErrorInvalidValue;
  const size_t lds = step2_lds_bytes(block_x, p.rows_per_tile, sizeof(T));
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  const unsigned tiles_y = static_cast<unsigned>((p.ny + p.rows_per_tile - 1) / p.rows_per_tile);
  const unsigned grid = static_cast<unsigned>(p.tiles_x) * tiles_y;
  if (p.store_nt)
    step2_kernel<T, kAuxNt><<<grid, block_x, lds, stream>>>(p);
  else
    step2_kernel<T, kAuxNone><<<grid, block_x, lds, stream>>>(p);
  return hipGetLastError();
}
template hipError_t launch_step2<float>(const StepParams<float>&, int, hipStream_t);
template hipError_t launch_step2<double>(const StepParams<double>&, int, hipStream_t);

}  // namespace lbmhip
