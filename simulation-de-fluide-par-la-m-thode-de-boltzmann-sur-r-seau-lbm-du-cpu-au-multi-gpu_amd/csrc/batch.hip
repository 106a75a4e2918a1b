// Batched persistent step over many independent domains of uneven sizes and parameters.
//
// All domains live in one pair of population allocations (each domain's 9 planes contiguous at
// its own offset), with one descriptor per domain (grid, offsets, omega, force prefactor,
// force) and a work-list of tiles (domain, column tile, row tile) in domain-major order. A fixed
// grid of G workgroups walks the list, in one of two orders picked by the autotuner: workgroup
// w takes tiles w, w + G, w + 2G, ... (the workgroups in flight sweep neighbouring tiles
// together), or the contiguous chunk [w * chunk, (w + 1) * chunk) with workgroups remapped so
// that each XCD takes a contiguous range; or, as a wave work-list, per-wavefront segments of
// 64 columns over a few rows. Each tile is updated with update_column, the same
// device code as the single-domain step, so each domain is bit-identical to a single-domain
// run. Per-domain parameters come from the descriptor table, not from process-global constants.
#include "batch.hpp"
#include "step_body.hpp"

namespace lbmhip {
namespace {

template <class T, int kAux, class PS>
__device__ __forceinline__ void batch_tile(const BatchParams<T>& p, int t) {
  const BatchTile tile = p.tiles[t];
  const BatchDomain<T> dm = p.domains[tile.domain];
  const DomainView<T> d{p.g + dm.offset, p.gn + dm.offset, p.links + dm.cell_offset,
                        p.segflags + dm.seg_offset, nullptr, dm.N, dm.nx, dm.ny, dm.nseg, 1,
                        dm.omega, dm.fp, dm.fx, dm.fy};
  const int x = tile.tx * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
  const int y0 = tile.ty * p.tile_rows;
  const int y1 = min(y0 + p.tile_rows, dm.ny);
  update_column<T, false, false, kAux, PS>(d, x, y0, y1, nullptr);
}

// Wave work-list: every wavefront takes 64-column segments of one domain (lanes beyond the
// domain's last column idle), so narrow domains do not leave most lanes of a wide tile empty.
// The descriptors of up to kBatchLdsDomains domains are staged in LDS once per workgroup.
template <class T, int kAux, class PS>
__device__ __forceinline__ void batch_waves(const BatchParams<T>& p) {
  __shared__ BatchDomain<T> lds[kBatchLdsDomains];
  const bool in_lds = p.ndomains <= kBatchLdsDomains;
  if (in_lds) {
    for (int k = static_cast<int>(threadIdx.x); k < p.ndomains; k += static_cast<int>(blockDim.x))
      lds[k] = p.domains[k];
    __syncthreads();
  }
  const int waves = static_cast<int>(blockDim.x) >> 6;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
  const int lane = static_cast<int>(threadIdx.x) & 63;
  const int stride = static_cast<int>(gridDim.x) * waves;
  for (int it = static_cast<int>(blockIdx.x) * waves + wave; it < p.ntiles; it += stride) {
    const BatchSeg sg = p.segs[it];
    const BatchDomain<T> dm = in_lds ? lds[sg.domain] : p.domains[sg.domain];
    const DomainView<T> d{p.g + dm.offset, p.gn + dm.offset, p.links + dm.cell_offset,
                          p.segflags + dm.seg_offset, nullptr, dm.N, dm.nx, dm.ny, dm.nseg, 1,
                          dm.omega, dm.fp, dm.fx, dm.fy};
    update_column<T, false, false, kAux, PS>(d, sg.seg * 64 + lane, sg.y0, sg.y1, nullptr);
  }
}

template <class T, int kAux, bool kOne>
__global__ void __launch_bounds__(1024) batch_kernel(const BatchParams<T> p) {
  using PS = typename std::conditional<kOne, PlanesOne, Planes>::type;
  if (p.order == 2) {
    batch_waves<T, kAux, PS>(p);
  } else if (p.order == 0) {
    for (int t = static_cast<int>(blockIdx.x); t < p.ntiles; t += static_cast<int>(gridDim.x))
      batch_tile<T, kAux, PS>(p, t);
  } else {
    const unsigned w = xcd_tile(blockIdx.x, gridDim.x);
    const int t0 = static_cast<int>(w) * p.chunk;
    const int t1 = min(t0 + p.chunk, p.ntiles);
    for (int t = t0; t < t1; ++t) batch_tile<T, kAux, PS>(p, t);
  }
}

}  // namespace

template <class T>
hipError_t launch_batch(const BatchParams<T>& p, int block_x, int workgroups, hipStream_t stream) {
  if (p.ntiles == 0 || workgroups <= 0) return hipSuccess;
  if (p.store_nt) {
    if (p.one_rsrc)
      batch_kernel<T, kAuxNt, true><<<workgroups, block_x, 0, stream>>>(p);
    else
      batch_kernel<T, kAuxNt, false><<<workgroups, block_x, 0, stream>>>(p);
  } else {
    if (p.one_rsrc)
      batch_kernel<T, kAuxNone, true><<<workgroups, block_x, 0, stream>>>(p);
    else
      batch_kernel<T, kAuxNone, false><<<workgroups, block_x, 0, stream>>>(p);
  }
  return hipGetLastError();
}
template hipError_t launch_batch<float>(const BatchParams<float>&, int, int, hipStream_t);
template hipError_t launch_batch<double>(const BatchParams<double>&, int, int, hipStream_t);

}  // namespace lbmhip
