// Descriptor table and work-list of the batched persistent step (batch.hip, batch.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace lbmhip {

template <class T>
struct BatchDomain {
  size_t offset;       // element offset of plane 0 in the population buffers
  size_t cell_offset;  // offset in the link-code array
  size_t seg_offset;   // offset in the segment-flag array
  size_t N;            // plane stride (cells rounded up to 64, plus kPlaneTail)
  int nx, ny, nseg, pad;
  T omega, fp, fx, fy;
};

struct BatchTile {
  int domain, tx, ty, pad;
};

// One wave's entry of a packed two-step work item: an item is blockDim / 64 consecutive
// entries, one per wave of the workgroup, holding up to one sub-tile per wave. A sub-tile (rows
// [y0, y0 + rows) of columns [x0, x0 + width) of one domain, y0 + rows <= ny) is walked by the
// width / 64 consecutive waves from thread lane0, with the LDS region that starts lds values
// (of the population type) into the workgroup's LDS. domain < 0: the wave has no sub-tile.
// nr_max: rows + 2 of the item's tallest sub-tile (the same in every entry of the item).
struct BatchWave {
  int domain, x0, y0, rows;
  int lane0, width, lds, nr_max;
};

// Work item of the wave work-list (order 2): one 64-column segment of one domain over rows
// [y0, y1), taken by one wavefront.
struct BatchSeg {
  int domain, seg, y0, y1;
};

// Domains whose descriptors the wave work-list kernel keeps in LDS (more: read from memory).
constexpr int kBatchLdsDomains = 256;

template <class T>
struct BatchParams {
  const T* g;
  T* gn;
  const uint16_t* links;
  const uint8_t* segflags;
  const BatchDomain<T>* domains;
  const BatchTile* tiles;
  const BatchWave* waves;  // the two-step list: ntiles items of blockDim / 64 entries
  const BatchSeg* segs;  // order 2
  int ndomains;
  int ntiles;            // tiles (orders 0, 1) or segments (order 2)
  int tile_rows;
  int chunk;     // order 1: tiles per workgroup
  int order;     // 0: workgroup w takes tiles w, w + G, w + 2G, ... (G = grid size);
                 // 1: a contiguous chunk of tiles, workgroups in per-XCD bands;
                 // 2: wave work-list, wave v of the grid takes segments v, v + V, ...
  int store_nt;  // 1: non-temporal stores
  int one_rsrc;  // 1: every domain's nine planes span < 4 GiB (one buffer resource each)
};

template <class T>
hipError_t launch_batch(const BatchParams<T>& p, int block_x, int workgroups, hipStream_t stream);

// Two steps per launch over the batch (step2.hip): p.waves / p.ntiles are the packed two-step
// work items (block_x / 64 entries each, LDS regions within batch_step2_lds_bytes(block_x,
// sizeof(T))), walked by `workgroups` persistent workgroups (0: one per item); every domain's
// plane stride holds kPlaneTail values past its cells. Bit-identical to two launch_batch
// steps.
template <class T>
hipError_t launch_batch_step2(const BatchParams<T>& p, int block_x, int workgroups,
                              hipStream_t stream);
// Workgroups of launch_batch_step2's kernel instance (stores non-temporal or not, one buffer
// resource or not) resident per CU with block_x threads, on the current device (0 on error).
template <class T>
int batch_step2_resident(int block_x, bool store_nt, bool one_rsrc);

}  // namespace lbmhip
