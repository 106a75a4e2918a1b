// C ABI of the plain solver (lbm_create ... lbm_autotune): the MI355X counterpart of SolverV3
// (src/v3_cuda_opt/solver_v3.hpp:27-73, solver_v3.cu:166-370).
//
// Ownership: the handle owns two population buffers, the link codes, the segment flags, an
// optional force field, a read-back staging buffer and a small reduction buffer, all in device
// memory, plus one HIP stream (default priority) and two events. All launches go to that
// stream; read-backs synchronise it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host_common.hpp"
#include "kernels.hpp"
#include "lbm_hip.h"
#include "solver_impl.hpp"

using namespace lbmhip;

namespace lbmhip {
const char* last_error();
}

namespace {

int check_handle(const lbm_solver* s) {
  return s ? LBM_OK : fail(LBM_EINVAL, "null solver handle");
}

}  // namespace

extern "C" {

const char* lbm_last_error(void) { return lbmhip::last_error(); }

int lbm_device_info(char* buf, size_t len) {
  int dev = 0;
  LBM_HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop{};
  LBM_HIP_TRY(hipGetDeviceProperties(&prop, dev));
  int rt = 0;
  LBM_HIP_TRY(hipRuntimeGetVersion(&rt));
  if (buf && len)
    std::snprintf(buf, len, "%s (%s), %d CUs, %.1f GB, HIP runtime %d", prop.name, prop.gcnArchName,
                  prop.multiProcessorCount, prop.totalGlobalMem / 1e9, rt);
  return LBM_OK;
}

int lbm_create(int nx, int ny, double tau, double force_x, double force_y, const uint8_t* solid,
               int precision, lbm_solver** out) {
  if (!out) return fail(LBM_EINVAL, "out handle pointer is null");
  *out = nullptr;
  if (nx < 1 || ny < 1) return fail(LBM_EINVAL, "grid too small");
  if (!solid) return fail(LBM_EINVAL, "solid mask size does not match the grid");
  if (precision != LBM_FLOAT && precision != LBM_DOUBLE && precision != LBM_FP16)
    return fail(LBM_EINVAL, "precision must be LBM_FLOAT, LBM_DOUBLE or LBM_FP16");
  auto* s = new lbm_solver();
  const int rc = s->create(nx, ny, tau, force_x, force_y, solid, precision);
  if (rc != LBM_OK) {
    delete s;
    return rc;
  }
  *out = s;
  return LBM_OK;
}

int lbm_destroy(lbm_solver* s) {
  if (!s) return LBM_OK;
  s->release();
  delete s;
  return LBM_OK;
}

int lbm_set_inflow_outflow(lbm_solver* s, double u_in) {
  LBM_TRY(check_handle(s));
  return s->set_inflow_outflow(u_in);
}

int lbm_init_equilibrium(lbm_solver* s, double rho0, double ux0, double uy0) {
  LBM_TRY(check_handle(s));
  return s->init_equilibrium(rho0, ux0, uy0);
}

int lbm_set_mask(lbm_solver* s, const uint8_t* solid) {
  LBM_TRY(check_handle(s));
  return s->set_mask(solid);
}

int lbm_set_force(lbm_solver* s, double fx, double fy) {
  LBM_TRY(check_handle(s));
  s->fx = fx;
  s->fy = fy;
  return LBM_OK;
}

int lbm_set_force_field(lbm_solver* s, const void* fxy) {
  LBM_TRY(check_handle(s));
  return s->set_force_field(fxy);
}

int lbm_set_tau(lbm_solver* s, double tau) {
  LBM_TRY(check_handle(s));
  return s->set_tau(tau);
}

int lbm_set_block(lbm_solver* s, int block_x, int rows_per_thread) {
  LBM_TRY(check_handle(s));
  LBM_TRY(check_block(block_x, rows_per_thread));
  s->block_x = block_x;
  s->rows = rows_per_thread;
  return LBM_OK;
}

int lbm_set_kernel_options(lbm_solver* s, int streaming_stores, int single_resource, int tile_order) {
  LBM_TRY(check_handle(s));
  if (tile_order < 0 || tile_order > 2) return fail(LBM_EINVAL, "tile_order must be 0, 1 or 2");
  s->store_nt = streaming_stores ? 1 : 0;
  s->one_rsrc = single_resource ? 1 : 0;
  s->xcd_bands = tile_order;
  return LBM_OK;
}

int lbm_get_kernel_options(const lbm_solver* s, int* streaming_stores, int* single_resource,
                           int* tile_order) {
  LBM_TRY(check_handle(s));
  if (streaming_stores) *streaming_stores = s->store_nt;
  if (single_resource) *single_resource = s->one_rsrc;
  if (tile_order) *tile_order = s->xcd_bands;
  return LBM_OK;
}

int lbm_get_block(const lbm_solver* s, int* block_x, int* rows_per_thread) {
  LBM_TRY(check_handle(s));
  if (block_x) *block_x = s->block_x;
  if (rows_per_thread) *rows_per_thread = s->rows;
  return LBM_OK;
}

int lbm_set_fp16_scale(lbm_solver* s, double scale) {
  LBM_TRY(check_handle(s));
  return s->set_fp16_scale(scale);
}

int lbm_step(lbm_solver* s) {
  LBM_TRY(check_handle(s));
  return s->step();
}

int lbm_run(lbm_solver* s, int steps) {
  LBM_TRY(check_handle(s));
  if (steps < 0) return fail(LBM_EINVAL, "steps must be >= 0");
  LBM_TRY(s->advance(steps));
  LBM_HIP_TRY(hipStreamSynchronize(s->stream));
  return LBM_OK;
}

int lbm_timed_run(lbm_solver* s, int steps, double* seconds) {
  LBM_TRY(check_handle(s));
  if (steps < 0) return fail(LBM_EINVAL, "steps must be >= 0");
  LBM_HIP_TRY(hipEventRecord(s->ev0, s->stream));
  LBM_TRY(s->advance(steps));
  LBM_HIP_TRY(hipEventRecord(s->ev1, s->stream));
  LBM_HIP_TRY(hipEventSynchronize(s->ev1));
  float ms = 0.0f;
  LBM_HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  if (seconds) *seconds = static_cast<double>(ms) * 1e-3;
  return LBM_OK;
}

int lbm_macroscopic(lbm_solver* s, void* rho, void* ux, void* uy) {
  LBM_TRY(check_handle(s));
  return s->macroscopic(rho, ux, uy);
}

int lbm_vorticity(lbm_solver* s, void* out, int periodic_x) {
  LBM_TRY(check_handle(s));
  return s->vorticity(out, periodic_x);
}

int lbm_total_mass(lbm_solver* s, double* mass) {
  LBM_TRY(check_handle(s));
  double r[4];
  LBM_TRY(s->reduce(r));
  if (mass) *mass = r[0] + r[1];
  return LBM_OK;
}

int lbm_momentum(lbm_solver* s, double* px, double* py) {
  LBM_TRY(check_handle(s));
  double r[4];
  LBM_TRY(s->reduce(r));
  if (px) *px = r[2];
  if (py) *py = r[3];
  return LBM_OK;
}

int lbm_populations(lbm_solver* s, void* out) {
  LBM_TRY(check_handle(s));
  return s->populations(out);
}

int lbm_populations_at(lbm_solver* s, int x, int y, void* out9) {
  LBM_TRY(check_handle(s));
  if (x < 0 || x >= s->nx || y < 0 || y >= s->ny) return fail(LBM_EINVAL, "cell out of the grid");
  return s->populations_at(x, y, out9);
}

int lbm_set_populations(lbm_solver* s, const void* in) {
  LBM_TRY(check_handle(s));
  return s->set_populations(in);
}

int lbm_get_state(lbm_solver* s, void* out) {
  LBM_TRY(check_handle(s));
  if (!out) return fail(LBM_EINVAL, "null output array");
  LBM_HIP_TRY(hipStreamSynchronize(s->stream));
  LBM_TRY(s->download_state(out, s->g));
  return LBM_OK;
}

int lbm_set_state(lbm_solver* s, const void* in) {
  LBM_TRY(check_handle(s));
  if (!in) return fail(LBM_EINVAL, "null input array");
  LBM_HIP_TRY(hipStreamSynchronize(s->stream));
  LBM_TRY(s->upload_state(s->g, in));
  LBM_TRY(s->upload_state(s->gn, in));
  return LBM_OK;
}

int lbm_set_steps_per_launch(lbm_solver* s, int steps_per_launch, int block_x, int rows,
                             int tile_order) {
  LBM_TRY(check_handle(s));
  if (tile_order < -1 || tile_order > 2) return fail(LBM_EINVAL, "tile_order must be -1 (keep), 0, 1 or 2");
  if (steps_per_launch != 1 && steps_per_launch != 2)
    return fail(LBM_EINVAL, "steps_per_launch must be 1 or 2");
  if (block_x != 0 && (block_x < 64 || block_x > 512 || block_x % 64))
    return fail(LBM_EINVAL, "block_x must be 0 or a multiple of 64 in 64..512");
  if (rows < -1) return fail(LBM_EINVAL, "rows must be -1 (automatic), 0 (keep) or positive");
  s->steps_per_launch = steps_per_launch;
  if (block_x && block_x != s->block2_x) {
    s->block2_x = block_x;
    s->rows2_auto = 0;
  }
  if (rows) s->rows2 = rows;
  if (tile_order >= 0) s->two_step_order = tile_order;
  return LBM_OK;
}

int lbm_get_steps_per_launch(const lbm_solver* s, int* steps_per_launch, int* block_x, int* rows,
                             int* tile_order, int* in_use) {
  LBM_TRY(check_handle(s));
  if (tile_order) *tile_order = s->two_step_order;
  if (steps_per_launch) *steps_per_launch = s->steps_per_launch;
  if (block_x) *block_x = s->block2_x;
  if (rows) *rows = s->two_step_rows();
  if (in_use) *in_use = s->two_step_eligible() ? 1 : 0;
  return LBM_OK;
}

int lbm_threads(const lbm_solver*) { return 0; }

int lbm_bench_info(const lbm_solver* s, char* buf, size_t len) {
  LBM_TRY(check_handle(s));
  int dev = 0;
  LBM_HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop{};
  LBM_HIP_TRY(hipGetDeviceProperties(&prop, dev));
  int rt = 0, drv = 0;
  LBM_HIP_TRY(hipRuntimeGetVersion(&rt));
  LBM_HIP_TRY(hipDriverGetVersion(&drv));
  static const char* names[] = {"float", "double", "fp16-storage"};
  // Some runtimes report an empty marketing name: fall back to the architecture.
  char arch[64] = {0};
  std::snprintf(arch, sizeof arch, "%s", prop.gcnArchName);
  if (char* colon = std::strchr(arch, ':')) *colon = 0;
  char gpu[300];
  std::snprintf(gpu, sizeof gpu, prop.name[0] ? "%s" : "AMD %s", prop.name[0] ? prop.name : arch);
  const bool two = s->two_step_eligible();  // run() advances pairs of steps per launch
  if (buf && len)
    std::snprintf(buf, len,
                  "\"gpu\": \"%s\", \"arch\": \"%s\", \"hip_runtime\": %d, \"hip_driver\": %d, "
                  "\"block\": \"%dx%d\", \"variant\": \"%s\", \"storage\": \"%s\", "
                  "\"stores\": \"%s\", \"planes\": \"%s\", \"tile_order\": \"%s\", "
                  "\"plane_stride\": %zu, \"steps_per_launch\": %d, \"two_step_tile\": \"%dx%d\", "
                  "\"two_step_order\": %d",
                  gpu, prop.gcnArchName, rt, drv, s->block_x, s->rows,
                  two ? "soa_pull_two_step" : "soa_pull_buffer", names[s->prec],
                  s->store_nt ? "nontemporal" : "default",
                  s->one_rsrc ? "one_resource" : "resource_per_plane",
                  s->xcd_bands == 1 ? "xcd_bands" : (s->xcd_bands == 2 ? "xcd_rows" : "round_robin"),
                  s->S, two ? 2 : 1, s->block2_x, s->two_step_rows(),
                  s->two_step_order);
  return LBM_OK;
}

int lbm_dims(const lbm_solver* s, int* nx, int* ny, int* precision) {
  LBM_TRY(check_handle(s));
  if (nx) *nx = s->nx;
  if (ny) *ny = s->ny;
  if (precision) *precision = s->prec;
  return LBM_OK;
}

int lbm_autotune(lbm_solver* s, int steps, int* best_block_x, int* best_rows, double* best_mlups) {
  LBM_TRY(check_handle(s));
  if (steps < 1) return fail(LBM_EINVAL, "steps must be >= 1");
  return s->autotune(steps, best_block_x, best_rows, best_mlups);
}

int lbm_momentum_exchange(lbm_solver* s, int x0, int y0, int w, int h, double* fx, double* fy) {
  LBM_TRY(check_handle(s));
  return s->momentum_exchange(x0, y0, w, h, fx, fy);
}

double lbm_poiseuille_force(int precision, int H, double tau, double u_max) {
  // real(8) * viscosity(tau) * u_max / (real(H) * real(H)), viscosity = (tau - 1/2) / 3,
  // all in the solver's real (src/common/geometry.hpp:24-26, d2q9.hpp:30).
  if (precision == LBM_DOUBLE) {
    const double nu = (tau - 0.5) / 3.0;
    return 8.0 * nu * u_max / (double(H) * double(H));
  }
  const float t = static_cast<float>(tau), u = static_cast<float>(u_max);
  const float nu = (t - 0.5f) / 3.0f;
  return static_cast<double>(8.0f * nu * u / (float(H) * float(H)));
}

int lbm_copy_ceiling(size_t bytes, int reps, double* copy_gbs, double* memcpy_gbs,
                     double* reference_gbs, double* nine_gbs) {
  return lbmhip::copy_ceiling(bytes, reps, copy_gbs, memcpy_gbs, reference_gbs, nine_gbs);
}

}  // extern "C"
