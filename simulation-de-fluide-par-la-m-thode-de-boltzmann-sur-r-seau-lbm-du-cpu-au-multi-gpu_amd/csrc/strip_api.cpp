// Strip engine: one GPU's share of the reference's v4 strip decomposition (src/v4_mpi_cuda/
// engine.hpp:36-61, engine_gpu.cu:27-131, strip_update.hpp:32-88).
//
// A strip holds ny_local owned rows plus ghost row 0 (below) and ghost row ny_local + 1 (above),
// SoA with N = nx * (ny_local + 2). The fused step kernel of step.hip runs on the owned rows with
// wrap_y = 0: the ghost rows are plain kernel inputs (they are never updated, and no y wrap is
// applied), exactly the v4 cell update. Halo contents: the ghost row below needs planes
// {2, 5, 6} (c_y = +1) of the top owned row of the strip below; the ghost row above needs planes
// {4, 7, 8} (c_y = -1) of the bottom owned row of the strip above; 3 * nx reals per message
// (engine.hpp:17-20). Two HIP streams (default priority) let the interior rows run while the
// boundary rows are packed, as in engine_gpu.cu:72-107.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "d2q9.hpp"
#include "host_common.hpp"
#include "kernels.hpp"
#include "lbm_hip.h"

using namespace lbmhip;

namespace {
constexpr int kHaloUp[3] = {2, 5, 6};    // c_y = +1, sent to the strip above
constexpr int kHaloDown[3] = {4, 7, 8};  // c_y = -1, sent to the strip below
}  // namespace

struct lbm_strip {
  int nx = 0, nyl = 0, prec = 0;
  double tau = 0.8, fx = 0.0, fy = 0.0, u_in = 0.0;
  bool open_x = false;
  int block_x = 512, rows = 1;  // the plain solver's fp32/fp64 defaults (LBM_BLOCK overrides)
  std::vector<uint8_t> solid;  // (nyl + 2) * nx, ghost rows included
  void* g = nullptr;
  void* gn = nullptr;
  uint16_t* links = nullptr;
  uint8_t* segflags = nullptr;
  int nseg = 0;
  void* halo = nullptr;  // pinned host staging, 12 * nx reals
  hipStream_t streams[2] = {nullptr, nullptr};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  size_t N() const { return static_cast<size_t>(nx) * (nyl + 2); }
  size_t rs() const { return real_size(prec); }
  size_t bytes() const { return N() * kQ * rs(); }
  char* plane_row(void* buf, int i, int y) const {
    return static_cast<char*>(buf) + (i * N() + static_cast<size_t>(y) * nx) * rs();
  }

  int create(int nx_, int nyl_, int prec_) {
    nx = nx_;
    nyl = nyl_;
    prec = prec_;
    LBM_TRY(parse_block_env(&block_x, &rows));
    for (auto& s : streams) LBM_HIP_TRY(hipStreamCreate(&s));
    LBM_HIP_TRY(hipEventCreate(&ev0));
    LBM_HIP_TRY(hipEventCreate(&ev1));
    if (hipMalloc(&g, bytes()) != hipSuccess || hipMalloc(&gn, bytes()) != hipSuccess ||
        hipMalloc(&links, N() * sizeof(uint16_t)) != hipSuccess)
      return fail(LBM_ENOMEM, "device allocation of the strip failed");
    nseg = (nx + 63) / 64;
    if (hipMalloc(&segflags, static_cast<size_t>(nyl + 2) * nseg) != hipSuccess)
      return fail(LBM_ENOMEM, "device allocation of the strip flags failed");
    LBM_HIP_TRY(hipMemset(g, 0, bytes()));
    LBM_HIP_TRY(hipMemset(gn, 0, bytes()));
    if (hipHostMalloc(&halo, 12 * static_cast<size_t>(nx) * rs(), hipHostMallocDefault) != hipSuccess)
      return fail(LBM_ENOMEM, "pinned host allocation of the halo buffers failed");
    solid.assign(N(), 0);
    return upload_links();
  }

  void release() {
    for (auto s : streams)
      if (s) (void)hipStreamSynchronize(s);
    for (void* p : {g, gn, static_cast<void*>(links), static_cast<void*>(segflags)})
      if (p) (void)hipFree(p);
    if (halo) (void)hipHostFree(halo);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (auto s : streams)
      if (s) (void)hipStreamDestroy(s);
  }

  int upload_links() {
    if (open_x && nx < 3) return fail(LBM_EINVAL, "an open domain needs nx >= 3");
    const Links L = build_links(nx, nyl + 2, solid.data(), open_x, false);
    for (auto s : streams) LBM_HIP_TRY(hipStreamSynchronize(s));
    LBM_HIP_TRY(hipMemcpy(links, L.code.data(), L.code.size() * sizeof(uint16_t),
                          hipMemcpyHostToDevice));
    LBM_HIP_TRY(hipMemcpy(segflags, L.seg.data(), L.seg.size(), hipMemcpyHostToDevice));
    return LBM_OK;
  }

  template <class T>
  StepParams<T> params(int first, int last) const {
    StepParams<T> p{};
    p.g = static_cast<const T*>(g);
    p.gn = static_cast<T*>(gn);
    p.links = links;
    p.segflags = segflags;
    p.force = nullptr;
    p.N = N();
    p.nx = nx;
    p.ny = nyl + 2;
    p.row0 = first;
    p.nrows = last - first + 1;
    p.rows_per_tile = rows;
    p.tiles_x = (nx + block_x - 1) / block_x;
    p.nseg = nseg;
    p.wrap_y = 0;
    p.one_rsrc = 1;  // used when the nine planes span < 4 GiB
    p.store_nt = 1;  // non-temporal stores (periodic-x, uniform-force kernel)
    p.xcd_bands = 0; // hardware round-robin workgroup order
    const T t = static_cast<T>(tau);
    p.omega = T(1) / t;
    p.fp = T(1) - T(0.5) * p.omega;
    p.fx = static_cast<T>(fx);
    p.fy = static_cast<T>(fy);
    T eq[kQ] = {};
    if (open_x) equilibrium_state<T>(T(1), static_cast<T>(u_in), T(0), eq);
    for (int i = 0; i < kQ; ++i) p.eq_in[i] = eq[i];
    return p;
  }

  int compute_rows(int first, int last, int s) {
    if (last < first) return LBM_OK;
    if (first < 1 || last > nyl || s < 0 || s > 1)
      return fail(LBM_EINVAL, "compute_rows: rows must lie in 1..ny_local, stream 0 or 1");
    const hipError_t e = prec == LBM_DOUBLE
                             ? launch_step(params<double>(first, last), open_x, block_x, streams[s])
                             : launch_step(params<float>(first, last), open_x, block_x, streams[s]);
    if (e != hipSuccess) return fail(LBM_EHIP, std::string("strip launch: ") + hipGetErrorString(e));
    return LBM_OK;
  }
};

extern "C" {

int lbm_strip_create(int nx, int ny_local, int precision, lbm_strip** out) {
  if (!out) return fail(LBM_EINVAL, "out handle pointer is null");
  *out = nullptr;
  if (nx < 1 || ny_local < 1) return fail(LBM_EINVAL, "strip too small");
  if (precision != LBM_FLOAT && precision != LBM_DOUBLE)
    return fail(LBM_EINVAL, "the strip engine runs in LBM_FLOAT or LBM_DOUBLE");
  auto* s = new lbm_strip();
  const int rc = s->create(nx, ny_local, precision);
  if (rc != LBM_OK) {
    s->release();
    delete s;
    return rc;
  }
  *out = s;
  return LBM_OK;
}

int lbm_strip_destroy(lbm_strip* s) {
  if (!s) return LBM_OK;
  s->release();
  delete s;
  return LBM_OK;
}

int lbm_strip_set_params(lbm_strip* s, double tau, double fx, double fy, int open_x, double u_in) {
  if (!s) return fail(LBM_EINVAL, "null strip handle");
  const bool ok = s->prec == LBM_DOUBLE ? tau > 0.5 : static_cast<float>(tau) > 0.5f;
  if (!ok) return fail(LBM_EINVAL, "tau must be > 0.5");
  s->tau = tau;
  s->fx = fx;
  s->fy = fy;
  s->u_in = u_in;
  if (static_cast<bool>(open_x) != s->open_x) {
    s->open_x = open_x != 0;
    return s->upload_links();
  }
  return LBM_OK;
}

int lbm_strip_upload_solid(lbm_strip* s, const uint8_t* solid_strip) {
  if (!s || !solid_strip) return fail(LBM_EINVAL, "null strip handle or mask");
  s->solid.assign(solid_strip, solid_strip + s->N());
  return s->upload_links();
}

int lbm_strip_upload_state(lbm_strip* s, const void* g) {
  if (!s || !g) return fail(LBM_EINVAL, "null strip handle or state");
  for (auto st : s->streams) LBM_HIP_TRY(hipStreamSynchronize(st));
  LBM_HIP_TRY(hipMemcpy(s->g, g, s->bytes(), hipMemcpyHostToDevice));
  LBM_HIP_TRY(hipMemcpy(s->gn, g, s->bytes(), hipMemcpyHostToDevice));
  return LBM_OK;
}

int lbm_strip_download_state(lbm_strip* s, void* g) {
  if (!s || !g) return fail(LBM_EINVAL, "null strip handle or state");
  for (auto st : s->streams) LBM_HIP_TRY(hipStreamSynchronize(st));
  LBM_HIP_TRY(hipMemcpy(g, s->g, s->bytes(), hipMemcpyDeviceToHost));
  return LBM_OK;
}

int lbm_strip_compute_rows(lbm_strip* s, int first, int last, int stream) {
  if (!s) return fail(LBM_EINVAL, "null strip handle");
  return s->compute_rows(first, last, stream);
}

int lbm_strip_pack_halo(lbm_strip* s, void* send_up, void* send_down) {
  if (!s || !send_up || !send_down) return fail(LBM_EINVAL, "null strip handle or buffer");
  const size_t row = static_cast<size_t>(s->nx) * s->rs();
  char* h = static_cast<char*>(s->halo);
  hipStream_t st = s->streams[0];  // same stream as the boundary rows
  for (int k = 0; k < 3; ++k) {
    LBM_HIP_TRY(hipMemcpyAsync(h + k * row, s->plane_row(s->gn, kHaloUp[k], s->nyl), row,
                               hipMemcpyDeviceToHost, st));
    LBM_HIP_TRY(hipMemcpyAsync(h + (3 + k) * row, s->plane_row(s->gn, kHaloDown[k], 1), row,
                               hipMemcpyDeviceToHost, st));
  }
  LBM_HIP_TRY(hipStreamSynchronize(st));  // the interior stream keeps running
  std::memcpy(send_up, h, 3 * row);
  std::memcpy(send_down, h + 3 * row, 3 * row);
  return LBM_OK;
}

int lbm_strip_unpack_halo(lbm_strip* s, const void* recv_from_down, const void* recv_from_up) {
  if (!s || !recv_from_down || !recv_from_up) return fail(LBM_EINVAL, "null strip handle or buffer");
  const size_t row = static_cast<size_t>(s->nx) * s->rs();
  char* h = static_cast<char*>(s->halo) + 6 * row;
  std::memcpy(h, recv_from_down, 3 * row);
  std::memcpy(h + 3 * row, recv_from_up, 3 * row);
  hipStream_t st = s->streams[0];
  for (int k = 0; k < 3; ++k) {
    LBM_HIP_TRY(hipMemcpyAsync(s->plane_row(s->gn, kHaloUp[k], 0), h + k * row, row,
                               hipMemcpyHostToDevice, st));
    LBM_HIP_TRY(hipMemcpyAsync(s->plane_row(s->gn, kHaloDown[k], s->nyl + 1), h + (3 + k) * row,
                               row, hipMemcpyHostToDevice, st));
  }
  return LBM_OK;
}

int lbm_strip_finish_step(lbm_strip* s) {
  if (!s) return fail(LBM_EINVAL, "null strip handle");
  for (auto st : s->streams) LBM_HIP_TRY(hipStreamSynchronize(st));
  std::swap(s->g, s->gn);
  return LBM_OK;
}

int lbm_strip_set_block(lbm_strip* s, int block_x, int rows_per_thread) {
  if (!s) return fail(LBM_EINVAL, "null strip handle");
  LBM_TRY(check_block(block_x, rows_per_thread));
  s->block_x = block_x;
  s->rows = rows_per_thread;
  return LBM_OK;
}

int lbm_strip_timed_steps(lbm_strip* s, int steps, double* seconds) {
  // The strip at its real shape with the ring closed on itself: after each update the ghost
  // rows of the new state are refreshed from the strip's own boundary rows by one device kernel
  // (the contents an RCCL/xGMI exchange would deliver), all on stream 0.
  if (!s || steps < 0) return fail(LBM_EINVAL, "null strip handle or negative steps");
  hipStream_t st = s->streams[0];
  for (auto x : s->streams) LBM_HIP_TRY(hipStreamSynchronize(x));
  LBM_HIP_TRY(hipEventRecord(s->ev0, st));
  for (int k = 0; k < steps; ++k) {
    const hipError_t e = s->prec == LBM_DOUBLE
                             ? launch_step(s->params<double>(1, s->nyl), s->open_x, s->block_x, st)
                             : launch_step(s->params<float>(1, s->nyl), s->open_x, s->block_x, st);
    if (e != hipSuccess) return fail(LBM_EHIP, std::string("strip launch: ") + hipGetErrorString(e));
    const hipError_t h = s->prec == LBM_DOUBLE
                             ? launch_halo_ring(static_cast<double*>(s->gn), s->N(), s->nx, s->nyl, st)
                             : launch_halo_ring(static_cast<float*>(s->gn), s->N(), s->nx, s->nyl, st);
    if (h != hipSuccess) return fail(LBM_EHIP, std::string("halo launch: ") + hipGetErrorString(h));
    std::swap(s->g, s->gn);
  }
  LBM_HIP_TRY(hipEventRecord(s->ev1, st));
  LBM_HIP_TRY(hipEventSynchronize(s->ev1));
  float ms = 0.0f;
  LBM_HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  if (seconds) *seconds = ms * 1e-3;
  return LBM_OK;
}

}  // extern "C"
