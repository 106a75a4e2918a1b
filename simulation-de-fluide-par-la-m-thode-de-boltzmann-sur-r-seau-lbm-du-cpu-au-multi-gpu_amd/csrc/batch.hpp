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
  size_t N;            // cells (plane stride)
  int nx, ny, nseg, pad;
  T omega, fp, fx, fy;
};

struct BatchTile {
  int domain, tx, ty, pad;
};

template <class T>
struct BatchParams {
  const T* g;
  T* gn;
  const uint16_t* links;
  const uint8_t* segflags;
  const BatchDomain<T>* domains;
  const BatchTile* tiles;
  int ntiles;
  int tile_rows;
  int chunk;  // tiles per workgroup
};

template <class T>
hipError_t launch_batch(const BatchParams<T>& p, int block_x, int workgroups, hipStream_t stream);

}  // namespace lbmhip
