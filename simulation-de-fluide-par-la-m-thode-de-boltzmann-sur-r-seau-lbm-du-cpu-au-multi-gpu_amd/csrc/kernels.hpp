// Kernel parameter blocks and host-side launchers of the D2Q9 step and its read-back kernels.
//
// Data layout in HBM (one domain):
//   g, gn     9 planes of N = nx * ny values (SoA), plane i at i * N, row-major cells
//             y * nx + x; g holds the post-collision pull state R (src/v3_cuda_opt/
//             solver_v3.hpp:23-26); the step reads g and writes gn.
//   links     N uint16 link codes (d2q9.hpp), read only in dirty segments.
//   segflags  one uint8 per (row, 64-column segment): 1 if any cell of the segment is solid,
//             pulls from a solid cell, or lies on an open-x edge. Clean segments (the bulk of
//             a channel) never read the link codes.
//   force     optional per-cell body force, 2 planes of N (fx, fy).
// Everything a launch needs is passed by value in these structs (kernel arguments), so two
// solvers never share parameters.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "d2q9.hpp"

namespace lbmhip {

template <class T>
struct StepParams {
  const T* g;
  T* gn;
  const uint16_t* links;
  const uint8_t* segflags;
  const T* force;      // nullptr: uniform (fx, fy)
  size_t N;            // plane stride
  int nx, ny;          // grid of the buffers (a strip includes its two ghost rows)
  int row0, nrows;     // rows updated: [row0, row0 + nrows)
  int rows_per_tile;   // rows walked by one thread
  int tiles_x;         // ceil(nx / blockDim.x)
  int nseg;            // ceil(nx / 64)
  int wrap_y;          // 1: periodic in y; 0: strip (rows row0-1 and row0+nrows exist)
  T omega, fp, fx, fy; // 1/tau, 1 - omega/2 (host-computed in T, solver_v3.cu:252-253)
  T eq_in[9];          // inlet equilibrium (open x)
  int store_nt;        // 1: non-temporal stores (periodic, uniform-force kernel only)
  int one_rsrc;        // 1: one buffer resource for the nine planes when they fit in 4 GiB
  int xcd_bands;       // tile order: 0 round-robin, 1 contiguous bands per XCD, 2 tile rows
                       // per XCD (row r on XCD r % 8; needs a multiple of 8 tile rows)
};

// Collision constants of the FP16-storage step (scaled units g' = S g), computed once per launch
// on the host (and per cell on the device for a per-cell force field). BGK towards the shifted
// equilibrium plus the Guo source; with b_k = omega w_k S rho for the weight class k of a
// direction (rest, axis, diagonal), cu = c_i . u and S drho = sum g':
//   omega geq'_i = b_k (4.5 cu^2 - 1.5 u.u) + 3 b_k cu + omega w_k S drho,
//   fp src'_i   = A_i + B_i cu - C_k u.F,  A_i = 3 S fp w_i c_i.F, B_i = 9 S fp w_i c_i.F,
//                                           C_k = 3 S fp w_k.
// A direction and its opposite (cu and -cu) share the even part and differ in the sign of the
// odd part 3 b_k cu + A_i. `unit` is 1: sums are written as fma(a, unit, b), exactly a + b, so
// that an operand converted from a stored binary16 value is read in its 16-bit form by a
// mixed-precision fused multiply-add instead of being converted by an instruction of its own
// (the compiler turns a literal fma(a, 1, b) back into an addition).
struct HalfCollide {
  float om1, invS, hfxS, hfyS, fx, fy, unit;  // hf.S = half the force in scaled units
  float ow[3], owS15[3], owS3[3], owS45[3];  // omega w_k and omega w_k S times 1.5, 3 and 4.5
  float A[9], B[9], C[3];

  __host__ __device__ void setup(float omega, float fp, float S, float inv_s, float fx_, float fy_) {
    om1 = 1.0f - omega;
    invS = inv_s;
    fx = fx_;
    fy = fy_;
    hfxS = 0.5f * fx_ * S;  // S a power of two: exact
    hfyS = 0.5f * fy_ * S;
    unit = 1.0f;
    for (int i = 0; i < kQ; ++i) {
      const float wi = weight<float>(i);
      const float cf = kCx[i] * fx_ + kCy[i] * fy_;
      A[i] = 3.0f * S * fp * wi * cf;
      B[i] = 9.0f * S * fp * wi * cf;
    }
    for (int k = 0; k < 3; ++k) {
      const float w = weight<float>(k == 0 ? 0 : (k == 1 ? 1 : 5));
      ow[k] = omega * w;
      owS15[k] = 1.5f * omega * w * S;
      owS3[k] = 3.0f * omega * w * S;
      owS45[k] = 4.5f * omega * w * S;
      C[k] = 3.0f * S * fp * w;
    }
  }
};

// FP16-storage step: halves hold g * scale; arithmetic in fp32 (fp16.hip).
struct StepParamsH {
  const uint16_t* g;
  uint16_t* gn;
  const uint16_t* links;
  const uint8_t* segflags;
  const float* force;
  size_t N;
  int nx, ny;
  int row0, nrows;
  int rows_per_tile;
  int tiles_x;
  int nseg;
  int wrap_y;
  float omega, fp, fx, fy;
  float scale, inv_scale;
  float eq_in[9];  // inlet equilibrium, already scaled
  uint32_t seed;   // stochastic-rounding seed of this step's pair (steps 2p, 2p + 1)
  uint32_t shift;  // 0 on the pair's even step, 16 on its odd step (step_random)
  int xcd_bands;   // tile order, as StepParams::xcd_bands
  HalfCollide col; // uniform-force collision constants
};

// Read-back parameters (gather of the post-streaming state), float / double storage.
template <class T>
struct GatherParams {
  const T* g;
  const uint16_t* links;
  const T* force;
  size_t N;
  int nx, ny;
  T fx, fy;
  T eq_in[9];
};

// Read-back parameters, FP16 storage (values returned unscaled, in float).
struct GatherParamsH {
  const uint16_t* g;
  const uint16_t* links;
  size_t N;
  int nx, ny;
  float eq_in_s[9];  // scaled inlet equilibrium
  float inv_scale;
};

// Launchers (step.hip, fp16.hip). All are asynchronous on `stream`.
template <class T>
hipError_t launch_step(const StepParams<T>& p, bool open_x, int block_x, hipStream_t stream);
hipError_t launch_step_fp16(const StepParamsH& p, bool open_x, int block_x, hipStream_t stream);

// out: 9N values, v0 AoS layout (cell * 9 + i), solid cells 0.
template <class T>
hipError_t launch_gather(const GatherParams<T>& p, T* out, hipStream_t stream);
hipError_t launch_gather_fp16(const GatherParamsH& p, float* out, hipStream_t stream);
template <class T>
hipError_t launch_gather_one(const GatherParams<T>& p, int x, int y, T* out9, hipStream_t stream);
hipError_t launch_gather_one_fp16(const GatherParamsH& p, int x, int y, float* out9,
                                  hipStream_t stream);
// rho, ux, uy (N each), solid cells 0.
template <class T>
hipError_t launch_macro(const GatherParams<T>& p, T* rho, T* ux, T* uy, hipStream_t stream);
hipError_t launch_macro_fp16(const GatherParamsH& p, float fx, float fy, const float* force,
                             float* rho, float* ux, float* uy, hipStream_t stream);
// FP64 sums over fluid cells of the post-streaming state: out[0] = number of fluid cells,
// out[1] = sum g, out[2] = sum cx g, out[3] = sum cy g. `partials` holds 4 doubles per block
// of the first pass (kReduceBlocks blocks).
constexpr int kReduceBlocks = 1024;
template <class T>
hipError_t launch_reduce(const GatherParams<T>& p, double* partials, double* out,
                         hipStream_t stream);
hipError_t launch_reduce_fp16(const GatherParamsH& p, double* partials, double* out,
                              hipStream_t stream);

// Vorticity of the velocity field ux, uy (device arrays of nx * ny) into w, 0 in solid cells
// (diag.hip; src/common/io.cpp:54-74); periodic_x = 0: one-sided differences at the open ends.
template <class T>
hipError_t launch_vorticity(const T* ux, const T* uy, const uint16_t* links, int nx, int ny,
                            int periodic_x, T* w, hipStream_t stream);

// Momentum-exchange force (FP64, out[0..1]) on the solids next to the fluid cells of the box
// [x0, x0 + w) x [y0, y0 + h) (diag.hip). S = float, double or uint16_t (FP16 storage).
template <class S>
hipError_t launch_exchange(const S* g, const uint16_t* links, size_t N, int nx, int x0, int y0,
                           int w, int h, float inv_scale, double* out, hipStream_t stream);

// Halo entries delivered by a two-step strip launch itself: as each of the sender's two
// outermost owned rows of the new state is stored, its message entries (halo_send) are also
// stored into the receiving strips' next-state buffers at their halo_recv rows (on one GPU a
// device store; strips on other GPUs would be reached through peer access). A solid cell's
// entries go to the receiver's discard slots instead, so solid ghost cells keep their values.
template <class T>
struct HaloOut {
  T* up;              // next-state buffer (plane 0) of the strip above: its ghost rows 0 and 1
  T* down;            // ... of the strip below: its ghost rows down_row and down_row + 1
  unsigned up_stride, down_stride;  // bytes between the receivers' planes
  unsigned up_tail, down_tail;      // byte offset of each receiver's discard slots in a plane
  int down_row;       // the receiver below's ny_local + 2
  int lo, hi;         // the sender's first and last owned rows (local)
};

// Values every population plane holds past its nx * ny cells (plane stride N >= nx * ny +
// kPlaneTail): launch_step2 stores the results of solid cells there instead of skipping them.
constexpr size_t kPlaneTail = 64;

// Two fused steps per launch (step2.hip): a periodic-y domain (periodic x, or open x with the
// inlet / outlet rules when open_x, nx >= 3), uniform force or a per-cell field (p.force), or
// the owned rows of a strip (periodic x, uniform force); planes < 4 GiB with the kPlaneTail
// tail; p.rows_per_tile = output rows per tile, block_x (64..512, a multiple of 64) columns per
// tile, p.tiles_x = ceil(nx / block_x). Bit-identical to two launches of launch_step (solid
// cells keep their values in both).
template <class T>
hipError_t launch_step2(const StepParams<T>& p, int block_x, hipStream_t stream,
                        bool open_x = false);
// The same over a strip's owned rows (wrap_y = 0), also storing its halo entries into the
// receivers (HaloOut). Each receiver's nine planes must span < 4 GiB (one buffer resource).
template <class T>
hipError_t launch_step2_halo(const StepParams<T>& p, const HaloOut<T>& h, int block_x,
                             hipStream_t stream);
// LDS bytes per workgroup of launch_step2 (independent of the tile height); elem = bytes of the
// values the walk carries (4 float, 8 double, 2 FP16 storage).
size_t step2_lds_bytes(int block_x, size_t elem);
// Values of one two-step walk's LDS region for a tile `width` threads wide (its ring and halo
// entries), and the LDS of a packed batch workgroup of block_x threads (every split into
// sub-tiles of whole waves fits).
size_t step2_region_elems(int width);
size_t batch_step2_lds_bytes(int block_x, size_t elem);
// Tile height for launch_step2: the tallest of 64, 48, 32, 24, 16 rows whose tile count fills
// the resident workgroup slots (occupancy of the instance p selects, on the current device)
// evenly with at least min_waves full waves (16 rows always qualifies), else the best-filling
// one.
template <class T>
int step2_tile_rows(const StepParams<T>& p, int block_x, int min_waves = 1);

// FP16 storage, two steps per launch (step2.hip): the conditions of launch_step2, and the launch
// starts on an even step (p.shift == 0; p.seed is the seed of the step pair). Bit-identical to
// two launch_step_fp16 steps. LDS per workgroup: step2_lds_bytes(block_x, 2) (the ring holds the
// step-1 values in binary16).
hipError_t launch_step2_fp16(const StepParamsH& p, int block_x, hipStream_t stream);
int step2_tile_rows_fp16(const StepParamsH& p, int block_x);

// ---- strips (one GPU's share of the v4 decomposition) ----------------------------------------
// A strip holds nyl owned rows (local rows ghost .. ghost + nyl - 1) and `ghost` ghost rows on
// each side. A halo message is a list of entries, each one row of nx values of one plane.
//   ghost = 1 (src/v4_mpi_cuda/engine.hpp:17-20): 3 entries. Up: planes {2, 5, 6} (c_y = +1) of
//     the top owned row, landing in ghost row 0 of the strip above; down: planes {4, 7, 8}
//     (c_y = -1) of the bottom owned row, landing in ghost row nyl + 1 of the strip below.
//   ghost = 2 (two steps per exchange): 12 entries. Up: the 9 planes of the top owned row (the
//     receiver's ghost row 1), then planes {2, 5, 6} of the row under it (its ghost row 0);
//     down: the 9 planes of the bottom owned row (the receiver's ghost row nyl + 2), then
//     planes {4, 7, 8} of the row above it (its ghost row nyl + 3). Two fused steps read a
//     strip's rows first - 2 .. last + 2: every plane of the inner ghost row (its own pulls and
//     bounce-back), and of the outer one only the planes that cross into the inner one.
__host__ __device__ constexpr int halo_entries(int ghost) { return ghost == 1 ? 3 : 12; }
// Entry k of the message a strip sends up (up = 1) or down: its plane and local row there.
__host__ __device__ inline void halo_send(int ghost, int nyl, int up, int k, int& plane, int& row) {
  const int moving = up ? (k == 0 ? 2 : (k == 1 ? 5 : 6)) : (k == 0 ? 4 : (k == 1 ? 7 : 8));
  if (ghost == 1) {
    plane = moving;
    row = up ? nyl : 1;
  } else if (k < kQ) {
    plane = k;
    row = up ? nyl + 1 : 2;
  } else {
    const int j = k - kQ;
    plane = up ? (j == 0 ? 2 : (j == 1 ? 5 : 6)) : (j == 0 ? 4 : (j == 1 ? 7 : 8));
    row = up ? nyl : 3;
  }
}
// Where entry k of a message received from the strip below (from_down = 1) or above lands.
__host__ __device__ inline void halo_recv(int ghost, int nyl, int from_down, int k, int& plane,
                                          int& row) {
  int sent_row = 0;
  halo_send(ghost, nyl, from_down, k, plane, sent_row);
  if (ghost == 1)
    row = from_down ? 0 : nyl + 1;
  else
    row = from_down ? (k < kQ ? 1 : 0) : (k < kQ ? nyl + 2 : nyl + 3);
}

// Strip ring closed on itself: the ghost rows of the state g (plane stride N, rows of nx
// values) refreshed from its own owned rows, as if the strips above and below were this one
// (the messages above, delivered on the device), one launch.
template <class T>
hipError_t launch_halo_ring(T* g, size_t N, int nx, int nyl, int ghost, hipStream_t stream);

// Grid-stride copy (HBM ceiling), n floats.
// variant 0: scalar (the reference's kernel), 1: 16-byte accesses, 2: 16-byte non-temporal.
hipError_t launch_copy(const float* a, float* b, size_t n, int grid, int block, int variant,
                       hipStream_t stream);
// Nine-array copy: b[k * n9 + i] = a[k * n9 + i] for k < 9, one i per thread (non-temporal
// when nt), over 9 * n9 floats.
hipError_t launch_copy9(const float* a, float* b, size_t n9, int nt, hipStream_t stream);

}  // namespace lbmhip
