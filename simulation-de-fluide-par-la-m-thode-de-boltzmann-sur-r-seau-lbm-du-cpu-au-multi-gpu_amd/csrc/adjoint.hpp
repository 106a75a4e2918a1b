// Parameters and launcher of the discrete-adjoint kernels (adjoint.hip).
//
// One launch runs up to one reverse step and up to three forward steps on disjoint buffers (a
// checkpointing schedule packs a reverse step with the recomputations that do not depend on
// it), each in its own blocks, one thread per cell, with a single memory round trip: the cells
// that can need a bounce-back operand (solid cells, links to solids) lie inside a few "dirty
// boxes" passed as kernel arguments, and only there the cell's own opposite populations are
// loaded, together with everything else, before the link code is looked at.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "kernels.hpp"

namespace lbmhip {

constexpr int kMaxBoxes = 4;
constexpr int kAdjMaxFwd = 3;

// Half-open boxes [x0, x1) x [y0, y1) covering every cell whose link code is not 0.
struct DirtyBoxes {
  int n;
  int x0[kMaxBoxes], x1[kMaxBoxes], y0[kMaxBoxes], y1[kMaxBoxes];
};

template <class T>
struct AdjGeom {
  const uint16_t* links;
  const T* force;  // per-cell force field: fx plane, fy plane (N each)
  size_t N;
  int nx, ny;
  T omega, fp;
  int px0, py0, pw, ph;  // probe window
  T k;                   // 1 / (window states * probe cells)
  double kd;             // the same in FP64 (J accumulation)
  DirtyBoxes boxes;
};

// One forward step src -> dst. jacc: FP64 accumulators of J, one per probe cell (row-major over
// the window), when the state src lies in the averaging window; null otherwise. With dst null
// the step only accumulates J (the last state of a run has no successor).
template <class T>
struct AdjFwd {
  const T* src;
  T* dst;
  double* jacc;
};

// One reverse step: mu_t from mu_{t+1} and R_t, plus the gradient accumulators.
template <class T>
struct AdjRev {
  const T* R;       // forward state R_t
  const T* mu_in;   // mu_{t+1}, or null for the last state (no later step)
  T* mu_out;        // mu_t
  T* dforce;        // dJ/dF accumulators (2N)
  double* tacc;     // dJ/dtau accumulators, one FP64 per cell
  double* jacc;     // J accumulators per probe cell when R_t lies in the window, else null
  int direct;       // 1 if R_t lies in the averaging window (direct term of J)
};

// Everything one launch needs, passed by value.
template <class T>
struct AdjArgs {
  AdjGeom<T> g;
  AdjRev<T> rev;
  AdjFwd<T> fwd[kAdjMaxFwd];
};

// Columns per block of the adjoint kernels (one thread per cell, one row per block).
constexpr int kAdjointThreads = 64;

// Blocks of one role (one step) of a launch.
__host__ __device__ inline unsigned adjoint_blocks(int nx, int ny) {
  return static_cast<unsigned>((nx + kAdjointThreads - 1) / kAdjointThreads) * static_cast<unsigned>(ny);
}

// A launch: the reverse step `rev` when has_rev, then nfwd forward steps fwd[0..nfwd).
template <class T>
hipError_t launch_adjoint(const AdjGeom<T>& g, const AdjRev<T>& rev, bool has_rev,
                          const AdjFwd<T>* fwd, int nfwd, hipStream_t stream);

// out[0] = sum of a[0..n) (fixed order: deterministic), on one block.
hipError_t launch_sum_f64(const double* a, size_t n, double* out, hipStream_t stream);

template <class T>
hipError_t launch_fill(T* a, size_t n, T v, hipStream_t stream);

}  // namespace lbmhip
