// Plain solver operations: allocation, host-side preparation (link codes, pull preimage),
// launches and read-backs. float and double share one template; FP16 storage has its own
// launchers (fp16.hip) behind the same state.
#include "solver_impl.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "d2q9.hpp"
#include "host_common.hpp"
#include "kernels.hpp"
#include "lbm_hip.h"

namespace lbmhip {

namespace {

template <class T>
StepParams<T> make_step_params(const SolverState& s) {
  StepParams<T> p{};
  p.g = static_cast<const T*>(s.g);
  p.gn = static_cast<T*>(s.gn);
  p.links = s.links;
  p.segflags = s.segflags;
  p.force = static_cast<const T*>(s.force);
  p.N = s.S;
  p.nx = s.nx;
  p.ny = s.ny;
  p.row0 = 0;
  p.nrows = s.ny;
  p.rows_per_tile = s.rows;
  p.tiles_x = (s.nx + s.block_x - 1) / s.block_x;
  p.nseg = s.nseg;
  p.wrap_y = 1;
  // omega and the force prefactor computed on the host in the solver real, as the reference
  // does once per step (src/v3_cuda_opt/solver_v3.cu:252-253).
  const T tau = static_cast<T>(s.tau);
  p.omega = T(1) / tau;
  p.fp = T(1) - T(0.5) * p.omega;
  p.fx = static_cast<T>(s.fx);
  p.fy = static_cast<T>(s.fy);
  T eq[kQ] = {};
  if (s.open_x) equilibrium_state<T>(T(1), static_cast<T>(s.u_in), T(0), eq);
  for (int i = 0; i < kQ; ++i) p.eq_in[i] = eq[i];
  p.store_nt = s.store_nt;
  p.one_rsrc = s.one_rsrc;
  p.xcd_bands = s.tile_order_for(s.rows);
  return p;
}

template <class T>
GatherParams<T> make_gather_params(const SolverState& s) {
  GatherParams<T> p{};
  p.g = static_cast<const T*>(s.g);
  p.links = s.links;
  p.force = static_cast<const T*>(s.force);
  p.N = s.S;
  p.nx = s.nx;
  p.ny = s.ny;
  p.fx = static_cast<T>(s.fx);
  p.fy = static_cast<T>(s.fy);
  T eq[kQ] = {};
  if (s.open_x) equilibrium_state<T>(T(1), static_cast<T>(s.u_in), T(0), eq);
  for (int i = 0; i < kQ; ++i) p.eq_in[i] = eq[i];
  return p;
}

GatherParamsH make_gather_params_h(const SolverState& s) {
  GatherParamsH p{};
  p.g = static_cast<const uint16_t*>(s.g);
  p.links = s.links;
  p.N = s.S;
  p.nx = s.nx;
  p.ny = s.ny;
  float eq[kQ] = {};
  if (s.open_x) equilibrium_state<float>(1.0f, static_cast<float>(s.u_in), 0.0f, eq);
  for (int i = 0; i < kQ; ++i) p.eq_in_s[i] = eq[i] * static_cast<float>(s.fp16_scale);
  p.inv_scale = static_cast<float>(1.0 / s.fp16_scale);
  return p;
}

StepParamsH make_step_params_h(const SolverState& s) {
  StepParamsH p{};
  p.g = static_cast<const uint16_t*>(s.g);
  p.gn = static_cast<uint16_t*>(s.gn);
  p.links = s.links;
  p.segflags = s.segflags;
  p.force = static_cast<const float*>(s.force);
  p.N = s.S;
  p.nx = s.nx;
  p.ny = s.ny;
  p.row0 = 0;
  p.nrows = s.ny;
  p.rows_per_tile = s.rows;
  p.tiles_x = (s.nx + s.block_x - 1) / s.block_x;
  p.nseg = s.nseg;
  p.wrap_y = 1;
  const float tau = static_cast<float>(s.tau);
  p.omega = 1.0f / tau;
  p.fp = 1.0f - 0.5f * p.omega;
  p.fx = static_cast<float>(s.fx);
  p.fy = static_cast<float>(s.fy);
  p.scale = static_cast<float>(s.fp16_scale);
  p.inv_scale = static_cast<float>(1.0 / s.fp16_scale);
  float eq[kQ] = {};
  if (s.open_x) equilibrium_state<float>(1.0f, static_cast<float>(s.u_in), 0.0f, eq);
  for (int i = 0; i < kQ; ++i) p.eq_in[i] = eq[i] * p.scale;
  p.seed = fp16_seed(s.step_count >> 1);
  p.shift = (s.step_count & 1u) ? 16u : 0u;
  p.xcd_bands = s.tile_order_for(s.rows);
  p.col.setup(p.omega, p.fp, p.scale, p.inv_scale, p.fx, p.fy);
  return p;
}

}  // namespace

// Stochastic-rounding seed of FP16 step pair `pair` (steps 2 pair and 2 pair + 1).
uint32_t fp16_seed(uint32_t pair) { return pair * 0x9E3779B9u + 0x7F4A7C15u; }

namespace {

// float -> scaled half, round to nearest even (host side, for uploads).
uint16_t to_half(float f) {
  _Float16 h = static_cast<_Float16>(f);
  uint16_t u;
  std::memcpy(&u, &h, 2);
  return u;
}

}  // namespace

size_t plane_stride(size_t N, size_t elem_bytes) {
  // At least kPlaneTail values past the last cell (the two-step kernel's discard slots).
  if (N * elem_bytes < (size_t(1) << 20)) return (N + 63) / 64 * 64 + kPlaneTail;
  const size_t q = 65536 / elem_bytes;
  return (N + q - 1) / q * q + 4352 / elem_bytes;
}

size_t SolverState::state_bytes() const { return S * kQ * storage_size(prec); }

int SolverState::upload_state(void* dev, const void* host) {
  const size_t es = storage_size(prec), row = cells() * es;
  for (int i = 0; i < kQ; ++i)
    LBM_HIP_TRY(hipMemcpy(static_cast<char*>(dev) + i * S * es,
                          static_cast<const char*>(host) + i * row, row, hipMemcpyHostToDevice));
  return LBM_OK;
}

int SolverState::download_state(void* host, const void* dev) {
  const size_t es = storage_size(prec), row = cells() * es;
  for (int i = 0; i < kQ; ++i)
    LBM_HIP_TRY(hipMemcpy(static_cast<char*>(host) + i * row,
                          static_cast<const char*>(dev) + i * S * es, row, hipMemcpyDeviceToHost));
  return LBM_OK;
}

int SolverState::create(int nx_, int ny_, double tau_, double fx_, double fy_, const uint8_t* solid_,
                        int prec_) {
  nx = nx_;
  ny = ny_;
  prec = prec_;
  fx = fx_;
  fy = fy_;
  LBM_TRY(set_tau(tau_));
  S = plane_stride(cells(), storage_size(prec));
  if (storage_size(prec) * S >= (size_t(1) << 32))
    return fail(LBM_EINVAL, "one population plane must stay below 4 GiB");
  solid.assign(solid_, solid_ + cells());
  // Launch defaults from shape scans of the 8192^2 channel on MI355X: one row per thread and
  // non-temporal stores for fp32/fp64 (512 x 1: more waves in flight, no per-wave row loop);
  // FP16 prefers 256 x 4 with ordinary stores. Tile order: XCD bands about +12 % for fp64;
  // fp32 within noise of round-robin for every order; FP16 fastest with XCD rows.
  // LBM_BLOCK overrides the shape.
  if (prec == LBM_FP16) {
    block_x = 256;
    rows = 4;
    store_nt = 0;
    xcd_bands = 2;  // XCD rows: 5.21 vs 4.90-5.17 TB/s round-robin (the 2-byte tiles share the
                    // most edge cache lines per byte)
  } else {
    block_x = 512;
    rows = 1;
    store_nt = 1;
    xcd_bands = prec == LBM_DOUBLE ? 1 : 0;
  }
  // two-step kernel: 512-column tiles in float; 256 in double (its LDS ring doubles; 512
  // columns exceed the LDS of a workgroup) and in FP16 storage (88 VGPRs: five 4-wave
  // workgroups per CU). 8192^2: fp32 512 x 64 146.3k MLUPS (256 x 64 128.5k), fp64 256 x 64
  // 69.6k (128 x 64 63.7k), FP16 256 x 64 221.6k (512 x 64 215.5k, 128 x 64 201.9k); the height
  // comes from step2_tile_rows(). Tile order: XCD bands, except FP16 (XCD rows: its 2-byte
  // tiles share the most edge lines per byte; 8192^2 230.8k against 229.1k with bands and
  // 227.2k round-robin).
  block2_x = prec == LBM_FLOAT ? 512 : 256;
  rows2 = -1;
  two_step_order = prec == LBM_FP16 ? 2 : 1;
  LBM_TRY(parse_block_env(&block_x, &rows));
  LBM_HIP_TRY(hipGetDevice(&device));
  LBM_HIP_TRY(hipStreamCreate(&stream));
  LBM_HIP_TRY(hipEventCreate(&ev0));
  LBM_HIP_TRY(hipEventCreate(&ev1));
  const size_t bytes = state_bytes();
  if (hipMalloc(&g, bytes) != hipSuccess || hipMalloc(&gn, bytes) != hipSuccess)
    return fail(LBM_ENOMEM, "device allocation of the population buffers failed");
  LBM_HIP_TRY(hipMemset(g, 0, bytes));
  LBM_HIP_TRY(hipMemset(gn, 0, bytes));
  if (hipMalloc(&links, cells() * sizeof(uint16_t)) != hipSuccess)
    return fail(LBM_ENOMEM, "device allocation of the link codes failed");
  nseg = (nx + 63) / 64;
  // whole dwords: the two-step kernel reads a flag with a scalar load of its aligned dword
  if (hipMalloc(&segflags, (static_cast<size_t>(ny) * nseg + 3) & ~size_t(3)) != hipSuccess)
    return fail(LBM_ENOMEM, "device allocation of the segment flags failed");
  if (hipMalloc(&red, sizeof(double) * (4 * kReduceBlocks + 4)) != hipSuccess)
    return fail(LBM_ENOMEM, "device allocation of the reduction buffer failed");
  return upload_links();
}

void SolverState::release() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (void* p : {g, gn, static_cast<void*>(links), static_cast<void*>(segflags), force, staging,
                  static_cast<void*>(red)})
    if (p) (void)hipFree(p);
  g = gn = force = staging = nullptr;
  links = nullptr;
  segflags = nullptr;
  red = nullptr;
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
  ev0 = ev1 = nullptr;
  stream = nullptr;
}

int SolverState::upload_links() {
  if (open_x && nx < 3) return fail(LBM_EINVAL, "an open domain needs nx >= 3");
  const Links L = build_links(nx, ny, solid.data(), open_x, true);
  LBM_HIP_TRY(hipStreamSynchronize(stream));
  LBM_HIP_TRY(hipMemcpy(links, L.code.data(), L.code.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  LBM_HIP_TRY(hipMemcpy(segflags, L.seg.data(), L.seg.size(), hipMemcpyHostToDevice));
  return LBM_OK;
}

int SolverState::set_inflow_outflow(double u) {
  open_x = true;
  u_in = u;
  return upload_links();
}

int SolverState::set_tau(double t) {
  const bool ok = prec == LBM_DOUBLE ? t > 0.5 : static_cast<float>(t) > 0.5f;
  if (!ok || !std::isfinite(t)) return fail(LBM_EINVAL, "tau must be > 0.5");
  tau = t;
  return LBM_OK;
}

int SolverState::set_fp16_scale(double scale) {
  if (!(scale > 0.0) || std::ldexp(1.0, std::ilogb(scale)) != scale || scale > 65536.0 * 16)
    return fail(LBM_EINVAL, "the FP16 scale must be a power of two");
  fp16_scale = scale;
  return LBM_OK;
}

int SolverState::set_mask(const uint8_t* s) {
  if (!s) return fail(LBM_EINVAL, "solid mask size does not match the grid");
  solid.assign(s, s + cells());
  return upload_links();
}

int SolverState::set_force_field(const void* fxy) {
  LBM_HIP_TRY(hipStreamSynchronize(stream));
  if (!fxy) {
    if (force) LBM_HIP_TRY(hipFree(force));
    force = nullptr;
    return LBM_OK;
  }
  // fx plane then fy plane, S values apart like the population planes
  const size_t rs = real_size(prec), row = cells() * rs;
  if (!force) {
    if (hipMalloc(&force, 2 * S * rs) != hipSuccess)
      return fail(LBM_ENOMEM, "device allocation of the force field failed");
    LBM_HIP_TRY(hipMemset(force, 0, 2 * S * rs));
  }
  LBM_HIP_TRY(hipMemcpy(force, fxy, row, hipMemcpyHostToDevice));
  LBM_HIP_TRY(hipMemcpy(static_cast<char*>(force) + S * rs, static_cast<const char*>(fxy) + row, row,
                        hipMemcpyHostToDevice));
  return LBM_OK;
}

int SolverState::init_equilibrium(double rho0, double ux0, double uy0) {
  const size_t n = cells() * kQ;
  auto upload = [&](const void* host) -> int {
    LBM_HIP_TRY(hipStreamSynchronize(stream));
    LBM_TRY(upload_state(g, host));
    return upload_state(gn, host);
  };
  if (prec == LBM_DOUBLE) {
    double eq[kQ];
    equilibrium_state<double>(rho0, ux0, uy0, eq);
    std::vector<double> R(n);
    preimage_uniform(nx, ny, solid.data(), open_x, eq, R.data());
    return upload(R.data());
  }
  float eq[kQ];
  equilibrium_state<float>(static_cast<float>(rho0), static_cast<float>(ux0), static_cast<float>(uy0),
                           eq);
  std::vector<float> R(n);
  preimage_uniform(nx, ny, solid.data(), open_x, eq, R.data());
  if (prec == LBM_FLOAT) return upload(R.data());
  std::vector<uint16_t> H(n);
  const float sc = static_cast<float>(fp16_scale);
  for (size_t k = 0; k < n; ++k) H[k] = to_half(R[k] * sc);
  return upload(H.data());
}

int SolverState::set_populations(const void* in) {
  if (!in) return fail(LBM_EINVAL, "null input array");
  const size_t n = cells() * kQ;
  auto upload = [&](const void* host) -> int {
    LBM_HIP_TRY(hipStreamSynchronize(stream));
    LBM_TRY(upload_state(g, host));
    return upload_state(gn, host);
  };
  if (prec == LBM_DOUBLE) {
    std::vector<double> R(n);
    preimage(nx, ny, solid.data(), open_x, static_cast<const double*>(in), R.data());
    return upload(R.data());
  }
  std::vector<float> R(n);
  preimage(nx, ny, solid.data(), open_x, static_cast<const float*>(in), R.data());
  if (prec == LBM_FLOAT) return upload(R.data());
  std::vector<uint16_t> H(n);
  const float sc = static_cast<float>(fp16_scale);
  for (size_t k = 0; k < n; ++k) H[k] = to_half(R[k] * sc);
  return upload(H.data());
}

int SolverState::step() {
  hipError_t e;
  if (prec == LBM_FLOAT)
    e = launch_step(make_step_params<float>(*this), open_x, block_x, stream);
  else if (prec == LBM_DOUBLE)
    e = launch_step(make_step_params<double>(*this), open_x, block_x, stream);
  else
    e = launch_step_fp16(make_step_params_h(*this), open_x, block_x, stream);
  if (e != hipSuccess) return fail(LBM_EHIP, std::string("step launch: ") + hipGetErrorString(e));
  std::swap(g, gn);
  ++step_count;
  return LBM_OK;
}

bool SolverState::two_step_eligible() const {
  // open x and force fields: float and double (FP16 storage keeps one step per launch there)
  if (steps_per_launch != 2 || ((open_x || force) && prec == LBM_FP16)) return false;
  if (S * storage_size(prec) >= (size_t(1) << 32)) return false;
  // the LDS ring holds the walk's values: binary16 for FP16 storage
  return step2_lds_bytes(block2_x, prec == LBM_DOUBLE ? 8 : (prec == LBM_FP16 ? 2 : 4)) <= 64 * 1024;
}

int SolverState::two_step_rows() const {
  if (rows2 > 0) return rows2;
  if (!rows2_auto) {
    if (prec == LBM_DOUBLE) {
      rows2_auto = step2_tile_rows(make_step_params<double>(*this), block2_x);
    } else if (prec == LBM_FP16) {
      rows2_auto = step2_tile_rows_fp16(make_step_params_h(*this), block2_x);
    } else {
      rows2_auto = step2_tile_rows(make_step_params<float>(*this), block2_x);
    }
  }
  return rows2_auto;
}

int SolverState::advance(int n) {
  // an FP16 two-step launch starts on an even step (its two steps share one hash per cell)
  if (prec == LBM_FP16 && (step_count & 1u) && n >= 2 && two_step_eligible()) {
    LBM_TRY(step());
    --n;
  }
  if (n >= 2 && two_step_eligible()) {
    hipError_t e = hipSuccess;
    const int rows_tile = two_step_rows();
    // XCD rows need whole groups of 8 tile rows; round-robin otherwise
    const int order = two_step_order == 2 && ((ny + rows_tile - 1) / rows_tile) % 8 ? 0 : two_step_order;
    for (; n >= 2 && e == hipSuccess; n -= 2) {
      if (prec == LBM_FP16) {
        StepParamsH p = make_step_params_h(*this);  // seed of the pair of step step_count
        p.rows_per_tile = rows_tile;
        p.tiles_x = (nx + block2_x - 1) / block2_x;
        p.xcd_bands = order;
        e = launch_step2_fp16(p, block2_x, stream);
      } else if (prec == LBM_FLOAT) {
        StepParams<float> p = make_step_params<float>(*this);
        p.rows_per_tile = rows_tile;
        p.tiles_x = (nx + block2_x - 1) / block2_x;
        p.xcd_bands = order;
        e = launch_step2(p, block2_x, stream, open_x);
      } else {
        StepParams<double> p = make_step_params<double>(*this);
        p.rows_per_tile = rows_tile;
        p.tiles_x = (nx + block2_x - 1) / block2_x;
        p.xcd_bands = order;
        e = launch_step2(p, block2_x, stream, open_x);
      }
      if (e == hipSuccess) {
        std::swap(g, gn);
        step_count += 2;
      }
    }
    if (e != hipSuccess) return fail(LBM_EHIP, std::string("two-step launch: ") + hipGetErrorString(e));
  }
  for (; n > 0; --n) LBM_TRY(step());
  return LBM_OK;
}

int SolverState::ensure_staging(size_t bytes) {
  if (staging_bytes >= bytes) return LBM_OK;
  if (staging) LBM_HIP_TRY(hipFree(staging));
  staging = nullptr;
  staging_bytes = 0;
  if (hipMalloc(&staging, bytes) != hipSuccess)
    return fail(LBM_ENOMEM, "device allocation of the read-back buffer failed");
  staging_bytes = bytes;
  return LBM_OK;
}

int SolverState::populations(void* out) {
  if (!out) return fail(LBM_EINVAL, "null output array");
  const size_t bytes = cells() * kQ * real_size(prec);
  LBM_TRY(ensure_staging(bytes));
  hipError_t e;
  if (prec == LBM_FLOAT)
    e = launch_gather(make_gather_params<float>(*this), static_cast<float*>(staging), stream);
  else if (prec == LBM_DOUBLE)
    e = launch_gather(make_gather_params<double>(*this), static_cast<double*>(staging), stream);
  else
    e = launch_gather_fp16(make_gather_params_h(*this), static_cast<float*>(staging), stream);
  if (e != hipSuccess) return fail(LBM_EHIP, std::string("gather launch: ") + hipGetErrorString(e));
  LBM_HIP_TRY(hipStreamSynchronize(stream));
  LBM_HIP_TRY(hipMemcpy(out, staging, bytes, hipMemcpyDeviceToHost));
  return LBM_OK;
}

int SolverState::populations_at(int x, int y, void* out9) {
  if (!out9) return fail(LBM_EINVAL, "null output array");
  const size_t bytes = kQ * real_size(prec);
  LBM_TRY(ensure_staging(bytes));
  hipError_t e;
  if (prec == LBM_FLOAT)
    e = launch_gather_one(make_gather_params<float>(*this), x, y, static_cast<float*>(staging), stream);
  else if (prec == LBM_DOUBLE)
    e = launch_gather_one(make_gather_params<double>(*this), x, y, static_cast<double*>(staging),
                          stream);
  else
    e = launch_gather_one_fp16(make_gather_params_h(*this), x, y, static_cast<float*>(staging),
                               stream);
  if (e != hipSuccess) return fail(LBM_EHIP, std::string("gather launch: ") + hipGetErrorString(e));
  LBM_HIP_TRY(hipStreamSynchronize(stream));
  LBM_HIP_TRY(hipMemcpy(out9, staging, bytes, hipMemcpyDeviceToHost));
  return LBM_OK;
}

int SolverState::macroscopic(void* rho, void* ux, void* uy) {
  if (!rho || !ux || !uy) return fail(LBM_EINVAL, "null output array");
  LBM_TRY(ensure_staging(3 * cells() * real_size(prec)));
  return macroscopic_into(staging, rho, ux, uy);
}

int SolverState::macroscopic_into(void* dev, void* rho, void* ux, void* uy) {
  if (!rho || !ux || !uy) return fail(LBM_EINVAL, "null output array");
  const size_t rs = real_size(prec), n = cells();
  char* st = static_cast<char*>(dev);
  hipError_t e;
  if (prec == LBM_FLOAT)
    e = launch_macro(make_gather_params<float>(*this), reinterpret_cast<float*>(st),
                     reinterpret_cast<float*>(st + n * rs), reinterpret_cast<float*>(st + 2 * n * rs),
                     stream);
  else if (prec == LBM_DOUBLE)
    e = launch_macro(make_gather_params<double>(*this), reinterpret_cast<double*>(st),
                     reinterpret_cast<double*>(st + n * rs),
                     reinterpret_cast<double*>(st + 2 * n * rs), stream);
  else {
    GatherParamsH p = make_gather_params_h(*this);
    e = launch_macro_fp16(p, static_cast<float>(fx), static_cast<float>(fy),
                          static_cast<const float*>(force), reinterpret_cast<float*>(st),
                          reinterpret_cast<float*>(st + n * rs),
                          reinterpret_cast<float*>(st + 2 * n * rs), stream);
  }
  if (e != hipSuccess) return fail(LBM_EHIP, std::string("macro launch: ") + hipGetErrorString(e));
  LBM_HIP_TRY(hipStreamSynchronize(stream));
  LBM_HIP_TRY(hipMemcpy(rho, st, n * rs, hipMemcpyDeviceToHost));
  LBM_HIP_TRY(hipMemcpy(ux, st + n * rs, n * rs, hipMemcpyDeviceToHost));
  LBM_HIP_TRY(hipMemcpy(uy, st + 2 * n * rs, n * rs, hipMemcpyDeviceToHost));
  return LBM_OK;
}

int SolverState::vorticity(void* out, int periodic_x) {
  if (!out) return fail(LBM_EINVAL, "null output array");
  const size_t rs = real_size(prec), n = cells();
  LBM_TRY(ensure_staging(4 * n * rs));
  char* st = static_cast<char*>(staging);
  hipError_t e;
  const int px = periodic_x < 0 ? (open_x ? 0 : 1) : (periodic_x ? 1 : 0);
  if (prec == LBM_DOUBLE) {
    double* f = reinterpret_cast<double*>(st);
    e = launch_macro(make_gather_params<double>(*this), f, f + n, f + 2 * n, stream);
    if (e == hipSuccess)
      e = launch_vorticity(f + n, f + 2 * n, links, nx, ny, px, f + 3 * n, stream);
  } else {
    float* f = reinterpret_cast<float*>(st);
    if (prec == LBM_FLOAT)
      e = launch_macro(make_gather_params<float>(*this), f, f + n, f + 2 * n, stream);
    else
      e = launch_macro_fp16(make_gather_params_h(*this), static_cast<float>(fx),
                            static_cast<float>(fy), static_cast<const float*>(force), f, f + n,
                            f + 2 * n, stream);
    if (e == hipSuccess)
      e = launch_vorticity(f + n, f + 2 * n, links, nx, ny, px, f + 3 * n, stream);
  }
  if (e != hipSuccess) return fail(LBM_EHIP, std::string("vorticity launch: ") + hipGetErrorString(e));
  LBM_HIP_TRY(hipStreamSynchronize(stream));
  LBM_HIP_TRY(hipMemcpy(out, st + 3 * n * rs, n * rs, hipMemcpyDeviceToHost));
  return LBM_OK;
}

int SolverState::reduce(double (&out)[4]) {
  hipError_t e;
  double* partials = red;
  double* result = red + 4 * kReduceBlocks;
  if (prec == LBM_FLOAT)
    e = launch_reduce(make_gather_params<float>(*this), partials, result, stream);
  else if (prec == LBM_DOUBLE)
    e = launch_reduce(make_gather_params<double>(*this), partials, result, stream);
  else
    e = launch_reduce_fp16(make_gather_params_h(*this), partials, result, stream);
  if (e != hipSuccess) return fail(LBM_EHIP, std::string("reduce launch: ") + hipGetErrorString(e));
  LBM_HIP_TRY(hipStreamSynchronize(stream));
  LBM_HIP_TRY(hipMemcpy(out, result, sizeof(out), hipMemcpyDeviceToHost));
  return LBM_OK;
}

int SolverState::momentum_exchange(int x0, int y0, int w, int h, double* fx, double* fy) {
  if (x0 < 0 || y0 < 0 || w < 1 || h < 1 || x0 + w > nx || y0 + h > ny)
    return fail(LBM_EINVAL, "momentum-exchange box outside the grid");
  double* out = red + 4 * kReduceBlocks;
  hipError_t e;
  if (prec == LBM_FLOAT)
    e = launch_exchange(static_cast<const float*>(g), links, S, nx, x0, y0, w, h, 1.0f, out,
                        stream);
  else if (prec == LBM_DOUBLE)
    e = launch_exchange(static_cast<const double*>(g), links, S, nx, x0, y0, w, h, 1.0f, out,
                        stream);
  else
    e = launch_exchange(static_cast<const uint16_t*>(g), links, S, nx, x0, y0, w, h,
                        static_cast<float>(1.0 / fp16_scale), out, stream);
  if (e != hipSuccess) return fail(LBM_EHIP, std::string("exchange launch: ") + hipGetErrorString(e));
  LBM_HIP_TRY(hipStreamSynchronize(stream));
  double r[2];
  LBM_HIP_TRY(hipMemcpy(r, out, sizeof r, hipMemcpyDeviceToHost));
  if (fx) *fx = r[0];
  if (fy) *fy = r[1];
  return LBM_OK;
}

int SolverState::autotune(int steps, int* best_bx, int* best_rows, double* best_mlups) {
  // Trials run on a copy of the state, which is restored afterwards whatever happens; when a
  // trial fails, the launch configuration is restored too. Every launch shape computes the
  // same bits (each cell's arithmetic does not depend on the shape).
  const size_t bytes = state_bytes();
  void* backup = nullptr;
  if (hipMalloc(&backup, bytes) != hipSuccess)
    return fail(LBM_ENOMEM, "device allocation of the autotune backup failed");
  const int k_bx = block_x, k_rows = rows, k_nt = store_nt, k_one = one_rsrc, k_band = xcd_bands,
            k_b2x = block2_x, k_r2 = rows2, k_r2a = rows2_auto, k_o2 = two_step_order;
  const uint32_t keep_count = step_count;  // FP16 rounding seeds stay as if no trial ran
  bool saved = false;
  int rc = hipStreamSynchronize(stream) == hipSuccess &&
                   hipMemcpy(backup, g, bytes, hipMemcpyDeviceToDevice) == hipSuccess
               ? LBM_OK
               : fail(LBM_EHIP, "autotune: saving the state failed");
  if (rc == LBM_OK) {
    saved = true;
    rc = autotune_trials(steps, best_bx, best_rows, best_mlups);
  }
  if (rc != LBM_OK) {
    block_x = k_bx;
    rows = k_rows;
    store_nt = k_nt;
    one_rsrc = k_one;
    xcd_bands = k_band;
    block2_x = k_b2x;
    rows2 = k_r2;
    rows2_auto = k_r2a;
    two_step_order = k_o2;
  }
  step_count = keep_count;
  bool restored = !saved;
  if (saved) {
    (void)hipGetLastError();  // a failed trial's launch error must not fail the restore
    restored = hipStreamSynchronize(stream) == hipSuccess &&
               hipMemcpy(g, backup, bytes, hipMemcpyDeviceToDevice) == hipSuccess &&
               hipMemcpy(gn, backup, bytes, hipMemcpyDeviceToDevice) == hipSuccess;
  }
  (void)hipFree(backup);
  if (rc != LBM_OK) return rc;
  return restored ? LBM_OK : fail(LBM_EHIP, "autotune: restoring the state failed");
}

int SolverState::autotune_trials(int steps, int* best_bx, int* best_rows, double* best_mlups) {
  const int keep_bx = block_x, keep_rows = rows;
  double best = -1.0;
  int bbx = keep_bx, brows = keep_rows, bnt = store_nt, bone = one_rsrc, bband = xcd_bands;
  for (int band : {0, 1, 2})
  for (int one : {1, 0})
  for (int nt : {0, 1})
  for (int bx : {128, 256, 512, 1024})
    for (int r : {1, 2, 4, 8, 16}) {
      block_x = bx;
      rows = r;
      store_nt = nt;
      one_rsrc = one;
      xcd_bands = band;
      int rc = step();  // warm-up (also checks the launch)
      if (rc != LBM_OK) continue;
      LBM_HIP_TRY(hipEventRecord(ev0, stream));
      for (int k = 0; k < steps; ++k) LBM_TRY(step());
      LBM_HIP_TRY(hipEventRecord(ev1, stream));
      LBM_HIP_TRY(hipEventSynchronize(ev1));
      float ms = 0.0f;
      LBM_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
      const double mlups = static_cast<double>(cells()) * steps / (ms * 1e-3) / 1e6;
      if (mlups > best) {
        best = mlups;
        bbx = bx;
        brows = r;
        bnt = nt;
        bone = one;
        bband = band;
      }
    }
  block_x = bbx;
  rows = brows;
  store_nt = bnt;
  one_rsrc = bone;
  xcd_bands = bband;
  // The two-step kernel, where run() takes it: tile columns, height (automatic, 32, 64) and
  // tile order; the reported rate is then the one of that path.
  if (two_step_eligible()) {
    const int k2x = block2_x, k2r = rows2, k2o = two_step_order;
    double best2 = -1.0;
    int b2x = k2x, b2r = k2r, b2o = k2o;
    for (int order : {0, 1, 2})
      for (int bx : {256, 512})
        for (int r : {-1, 32, 64}) {
          block2_x = bx;
          rows2 = r;
          rows2_auto = 0;
          two_step_order = order;
          if (!two_step_eligible()) continue;
          if (advance(2) != LBM_OK) continue;  // warm-up (also checks the launch)
          LBM_HIP_TRY(hipEventRecord(ev0, stream));
          LBM_TRY(advance(2 * ((steps + 1) / 2)));
          LBM_HIP_TRY(hipEventRecord(ev1, stream));
          LBM_HIP_TRY(hipEventSynchronize(ev1));
          float ms = 0.0f;
          LBM_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
          const double mlups =
              static_cast<double>(cells()) * 2 * ((steps + 1) / 2) / (ms * 1e-3) / 1e6;
          if (mlups > best2) {
            best2 = mlups;
            b2x = bx;
            b2r = r;
            b2o = order;
          }
        }
    block2_x = b2x;
    rows2 = b2r;
    rows2_auto = 0;
    two_step_order = b2o;
    if (best2 > 0.0) best = best2;
  }
  if (best_bx) *best_bx = bbx;
  if (best_rows) *best_rows = brows;
  if (best_mlups) *best_mlups = best;
  return LBM_OK;
}

int copy_ceiling(size_t bytes, int reps, double* copy_gbs, double* memcpy_gbs,
                 double* reference_gbs, double* nine_gbs) {
  if (bytes < 4 || reps < 1) return fail(LBM_EINVAL, "need bytes >= 4 and reps >= 1");
  const size_t n = bytes / sizeof(float);
  float *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * sizeof(float)) != hipSuccess) return fail(LBM_ENOMEM, "copy ceiling alloc");
  if (hipMalloc(&b, n * sizeof(float)) != hipSuccess) {
    (void)hipFree(a);
    return fail(LBM_ENOMEM, "copy ceiling alloc");
  }
  hipStream_t st;
  hipEvent_t e0, e1;
  LBM_HIP_TRY(hipStreamCreate(&st));
  LBM_HIP_TRY(hipEventCreate(&e0));
  LBM_HIP_TRY(hipEventCreate(&e1));
  LBM_HIP_TRY(hipMemsetAsync(a, 0, n * sizeof(float), st));
  LBM_HIP_TRY(hipMemsetAsync(b, 0, n * sizeof(float), st));
  int dev = 0, cus = 0;
  LBM_HIP_TRY(hipGetDevice(&dev));
  LBM_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  auto best_ms = [&](auto&& launch) -> double {
    launch();
    (void)hipStreamSynchronize(st);
    double best = 1e30;
    for (int r = 0; r < reps; ++r) {
      (void)hipEventRecord(e0, st);
      launch();
      (void)hipEventRecord(e1, st);
      (void)hipEventSynchronize(e1);
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      best = std::min(best, static_cast<double>(ms));
    }
    return best;
  };
  // The reference's scalar kernel (grid = CUs x 32, block 256) and 16-byte variants over grid
  // sizes up to about one 16-byte access per lane (the largest grids keep the most bytes in
  // flight); the ceiling is the best of all (a higher ceiling is the conservative choice).
  double kms = 1e30;
  for (int variant : {0, 1, 2})
    for (int per_cu : {8, 32, 128, 512, 2048}) {
      kms = std::min(kms, best_ms([&] { (void)launch_copy(a, b, n, cus * per_cu, 256, variant, st); }));
      LBM_HIP_TRY(hipGetLastError());
    }
  const double rms = best_ms([&] { (void)launch_copy(a, b, n, cus * 32, 256, 0, st); });
  double nms = 1e30;
  const size_t n9 = n / kQ;
  for (int nt : {0, 1}) nms = std::min(nms, best_ms([&] { (void)launch_copy9(a, b, n9, nt, st); }));
  LBM_HIP_TRY(hipGetLastError());
  const double mms = best_ms([&] {
    (void)hipMemcpyAsync(b, a, n * sizeof(float), hipMemcpyDeviceToDevice, st);
  });
  LBM_HIP_TRY(hipStreamSynchronize(st));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  (void)hipFree(a);
  (void)hipFree(b);
  const double total = 2.0 * static_cast<double>(n * sizeof(float));
  if (copy_gbs) *copy_gbs = total / (kms * 1e-3) / 1e9;
  if (memcpy_gbs) *memcpy_gbs = total / (mms * 1e-3) / 1e9;
  if (reference_gbs) *reference_gbs = total / (rms * 1e-3) / 1e9;
  if (nine_gbs) *nine_gbs = 2.0 * static_cast<double>(kQ * n9 * sizeof(float)) / (nms * 1e-3) / 1e9;
  return LBM_OK;
}

}  // namespace lbmhip
