// Batched persistent step over many independent domains of uneven sizes and parameters.
//
// All domains live in one pair of population allocations (each domain's 9 planes contiguous at
// its own offset), with one descriptor per domain (grid, offsets, omega, force prefactor,
// force) and a work-list of tiles (domain, column tile, row tile) in domain-major order. A fixed
// grid of workgroups walks the list: workgroup w (remapped so that each XCD takes a contiguous
// range) owns the contiguous chunk [w * chunk, (w + 1) * chunk) of tiles and updates each with
// update_column, the same device code as the single-domain step, so each domain is bit-identical
// to a single-domain run. Per-domain parameters come from the descriptor table, not from
// process-global constants.
#include "batch.hpp"
#include "step_body.hpp"

namespace lbmhip {
namespace {

template <class T>
__global__ void __launch_bounds__(1024) batch_kernel(const BatchParams<T> p) {
  const unsigned w = xcd_tile(blockIdx.x, gridDim.x);
  const int t0 = static_cast<int>(w) * p.chunk;
  const int t1 = min(t0 + p.chunk, p.ntiles);
  for (int t = t0; t < t1; ++t) {
    const BatchTile tile = p.tiles[t];
    const BatchDomain<T> dm = p.domains[tile.domain];
    const DomainView<T> d{p.g + dm.offset, p.gn + dm.offset, p.links + dm.cell_offset,
                          p.segflags + dm.seg_offset, nullptr, dm.N, dm.nx, dm.ny, dm.nseg, 1,
                          dm.omega, dm.fp, dm.fx, dm.fy};
    const int x = tile.tx * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    const int y0 = tile.ty * p.tile_rows;
    const int y1 = min(y0 + p.tile_rows, dm.ny);
    update_column<T, false, false>(d, x, y0, y1, nullptr);
  }
}

}  // namespace

template <class T>
hipError_t launch_batch(const BatchParams<T>& p, int block_x, int workgroups, hipStream_t stream) {
  if (p.ntiles == 0 || workgroups <= 0) return hipSuccess;
  batch_kernel<T><<<workgroups, block_x, 0, stream>>>(p);
  return hipGetLastError();
}
template hipError_t launch_batch<float>(const BatchParams<float>&, int, int, hipStream_t);
template hipError_t launch_batch<double>(const BatchParams<double>&, int, int, hipStream_t);

}  // namespace lbmhip
