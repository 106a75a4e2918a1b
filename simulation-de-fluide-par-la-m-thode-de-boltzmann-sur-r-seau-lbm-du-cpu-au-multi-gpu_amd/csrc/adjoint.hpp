// Parameters and launchers of the discrete-adjoint kernels (adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "kernels.hpp"

namespace lbmhip {

template <class T>
struct AdjointParams {
  const T* R;          // forward state R_t (pull layout)
  const T* mu_in;      // mu_{t+1}, or null for the last state (no later step)
  T* mu_out;           // mu_t
  const uint16_t* links;
  const T* force;      // per-cell force field (fx plane, fy plane)
  T* dforce;           // dJ/dF accumulators (2N)
  double* partials;    // two slots per block: dJ/dtau, J
  size_t N;
  int nx, ny;
  T omega, fp;
  int direct;          // 1 if state t lies in the averaging window
  int px0, py0, pw, ph;
  T k;                 // 1 / (window steps * probe cells)
  double kd;           // the same in FP64, for the J partials
};

// Block size of the reverse kernels, and tile width of the forward part of a fused launch.
constexpr int kAdjointThreads = 256;

__host__ __device__ inline unsigned adjoint_blocks(size_t N) {
  return static_cast<unsigned>((N + kAdjointThreads - 1) / kAdjointThreads);
}

template <class T>
hipError_t launch_adjoint_step(const AdjointParams<T>& p, hipStream_t stream);
template <class T>
hipError_t launch_adjoint_fused(const AdjointParams<T>& a, const StepParams<T>& f,
                                hipStream_t stream);
template <class T>
hipError_t launch_probe(const GatherParams<T>& p, int px0, int py0, int pw, int ph, double k,
                        double* acc, hipStream_t stream);
template <class T>
hipError_t launch_fill(T* a, size_t n, T v, hipStream_t stream);

}  // namespace lbmhip
