// Host-side helpers shared by the C-ABI translation units: error state, HIP status checks,
// link-code construction and the pull preimage of a post-streaming state.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "d2q9.hpp"
#include "lbm_hip.h"

namespace lbmhip {

// Records the message of the last failure on this thread and returns `code`.
int fail(int code, const std::string& msg);

#define LBM_HIP_TRY(expr)                                                                   \
  do {                                                                                      \
    const hipError_t lbm_e_ = (expr);                                                       \
    if (lbm_e_ != hipSuccess)                                                               \
      return ::lbmhip::fail(LBM_EHIP, std::string(#expr) + ": " + hipGetErrorString(lbm_e_)); \
  } while (0)

#define LBM_TRY(expr)          \
  do {                         \
    const int lbm_r_ = (expr); \
    if (lbm_r_ != LBM_OK) return lbm_r_; \
  } while (0)

inline size_t storage_size(int precision) {
  return precision == LBM_DOUBLE ? 8 : (precision == LBM_FP16 ? 2 : 4);
}
inline size_t real_size(int precision) { return precision == LBM_DOUBLE ? 8 : 4; }

// Per-cell link codes (d2q9.hpp) and per-(row, 64-column segment) dirty flags.
//   wrap_y = true: periodic in y (a whole domain); false: a strip whose first and last rows are
//   ghost rows (they get only the solid bit and are never updated).
struct Links {
  std::vector<uint16_t> code;
  std::vector<uint8_t> seg;
  int nseg = 0;
};
Links build_links(int nx, int ny, const uint8_t* solid, bool open_x, bool wrap_y);

// Pull preimage R (SoA, plane i at i * N) of a post-streaming state S (v0 AoS layout): slot
// (c, i) of R streams to (c + e_i, i), or to (c, opp i) when c + e_i is solid; a slot leaving
// an open domain is never read and is set to 0; solid cells are 0.
template <class T>
void preimage(int nx, int ny, const uint8_t* solid, bool open_x, const T* S, T* R);

// Preimage of the uniform equilibrium (rho0, ux0, uy0), as SolverV3::init_equilibrium builds it
// (src/v3_cuda_opt/solver_v3.cu:194-226): leaving slots of an open domain get eq[i].
template <class T>
void preimage_uniform(int nx, int ny, const uint8_t* solid, bool open_x, const T* eq, T* R);

// Parses "BXxBY" (the reference's LBM_BLOCK syntax, solver_v3.cu:145-153); 0 if unset.
int parse_block_env(int* bx, int* by);
int check_block(int bx, int rows);

}  // namespace lbmhip
