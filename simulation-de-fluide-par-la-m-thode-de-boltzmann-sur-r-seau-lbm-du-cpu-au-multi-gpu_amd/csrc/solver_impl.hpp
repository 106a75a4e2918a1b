// State and operations of one plain solver handle (see solver.cpp for the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace lbmhip {

struct SolverState {
  int nx = 0, ny = 0, prec = 0;
  int device = 0;  // the HIP device current at creation: it holds every buffer of the handle
  double tau = 0.0, fx = 0.0, fy = 0.0;  // as given; converted to the solver real at use
  std::vector<uint8_t> solid;
  bool open_x = false;
  double u_in = 0.0;
  int block_x = 256, rows = 8;  // launch shape (tile = block_x columns x rows)
  int store_nt = 0;             // non-temporal stores in the periodic uniform-force step
  int one_rsrc = 1;             // one buffer resource for all planes (when < 4 GiB)
  int xcd_bands = 0;            // tile order: 0 round-robin, 1 XCD bands, 2 XCD rows
  // The tile order a launch uses: XCD rows need a multiple of 8 tile rows, else round-robin.
  int tile_order_for(int rows_per_tile) const {
    const int tiles_y = (ny + rows_per_tile - 1) / rows_per_tile;
    return xcd_bands == 2 && tiles_y % 8 != 0 ? 0 : xcd_bands;
  }
  int steps_per_launch = 2;     // run(n) with two fused steps per launch where eligible
  int block2_x = 512;  // tile columns of the two-step kernel
  int rows2 = -1;      // tile rows; -1: chosen by step2_tile_rows() (64 on large grids)
  mutable int rows2_auto = 0;  // that height, once computed (0: not yet)
  // Tile order of the two-step kernel: XCD bands (a tile's neighbours above, below and beside it
  // mostly run on its own XCD, whose L2 then serves the halo rows and columns): 135.2k MLUPS at
  // 8192^2 fp32 against 132.0k round-robin and 131.6k XCD rows.
  int two_step_order = 1;
  double fp16_scale = 4096.0;   // stored half = g * scale (LBM_FP16)
  uint32_t step_count = 0;      // steps taken (seeds the FP16 stochastic rounding)

  void* g = nullptr;   // pull state R, SoA, storage type
  void* gn = nullptr;  // next state
  uint16_t* links = nullptr;
  uint8_t* segflags = nullptr;
  int nseg = 0;
  void* force = nullptr;     // per-cell force field (2N reals), or null
  void* staging = nullptr;   // read-back staging (9N reals)
  double* red = nullptr;     // reduction partials + result
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // Population planes are S >= nx * ny + kPlaneTail values apart (plane_stride(); large grids
  // are padded further to spread the planes over the memory channels).
  size_t S = 0;
  size_t cells() const { return static_cast<size_t>(nx) * ny; }
  size_t state_bytes() const;  // 9 planes of S values of the storage type
  // Raw state in host layout (9 planes of nx * ny values, contiguous) to / from a device buffer.
  int upload_state(void* dev, const void* host);
  int download_state(void* host, const void* dev);

  int create(int nx, int ny, double tau, double fx, double fy, const uint8_t* solid, int prec);
  void release();
  int set_inflow_outflow(double u);
  int init_equilibrium(double rho0, double ux0, double uy0);
  int set_mask(const uint8_t* solid);
  int set_force_field(const void* fxy);
  int set_tau(double tau);
  int set_fp16_scale(double scale);
  int step();
  // n steps: pairs of steps in one launch of the two-step kernel where eligible (periodic,
  // uniform force, float or double with the nine planes < 4 GiB), the rest single steps.
  int advance(int n);
  bool two_step_eligible() const;
  int two_step_rows() const;  // tile rows the two-step kernel uses
  int vorticity(void* out, int periodic_x);  // periodic_x < 0: the solver's own x boundary
  int populations(void* out);
  int populations_at(int x, int y, void* out9);
  int set_populations(const void* in);
  int macroscopic(void* rho, void* ux, void* uy);
  // The same into host arrays, through `dev`: a device buffer of at least 3 * nx * ny reals
  // that receives the fields before the copies (the handle's staging buffer, or a caller's).
  int macroscopic_into(void* dev, void* rho, void* ux, void* uy);
  int reduce(double (&out)[4]);
  int autotune(int steps, int* bx, int* rows, double* mlups);
  int momentum_exchange(int x0, int y0, int w, int h, double* fx, double* fy);

 private:
  int autotune_trials(int steps, int* bx, int* rows, double* mlups);  // autotune's timed trials
  int upload_links();
  int ensure_staging(size_t bytes);
  size_t staging_bytes = 0;
};

// Plane stride (values) of the population buffers of an N-cell grid: planes of 1 MiB and more
// are padded so that plane i starts i * 4352 bytes (mod 64 KiB) after plane 0 instead of at the
// same offset: the nine streams of a cell then spread over DRAM channels and banks (measured
// 1-2 % faster at 8192^2 on MI355X). Smaller grids keep S = N.
size_t plane_stride(size_t N, size_t elem_bytes);
uint32_t fp16_seed(uint32_t pair);  // stochastic-rounding seed of FP16 step pair `pair`

int copy_ceiling(size_t bytes, int reps, double* copy_gbs, double* memcpy_gbs,
                 double* reference_gbs, double* nine_gbs);

}  // namespace lbmhip

// The opaque handle of the C ABI (include/lbm_hip.h) is the solver state itself.
struct lbm_solver : lbmhip::SolverState {};
