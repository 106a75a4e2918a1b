// Strip engine: one GPU's share of the reference's v4 strip decomposition (src/v4_mpi_cuda/
// engine.hpp:36-61, engine_gpu.cu:27-131, strip_update.hpp:32-88).
//
// A strip holds ny_local owned rows plus `ghost` ghost rows below and above; the state is SoA
// with nx * (ny_local + 2 * ghost) cells per plane (host layout contiguous; on the device the
// planes are plane_stride() values apart and carry the two-step kernel's discard slots).
//   ghost = 1 is the reference's strip: the fused step kernel of step.hip runs on the owned
//     rows with wrap_y = 0, the ghost rows are plain kernel inputs (never updated, no y wrap),
//     exactly the v4 cell update; halo messages carry planes {2, 5, 6} up and {4, 7, 8} down,
//     3 * nx values each (engine.hpp:17-20).
//   ghost = 2 adds two fused steps per launch (step2.hip, the walk of the whole-domain
//     two-step kernel restricted to the owned rows): a launch reads rows first - 2 ..
//     last + 2, so one exchange of 12 * nx values per direction (kernels.hpp halo_send) serves
//     two steps. Single steps work too (the exchange refreshes both ghost rows).
// Two HIP streams (default priority) let the interior rows run while the boundary rows are
// packed, as in engine_gpu.cu:72-107.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "d2q9.hpp"
#include "host_common.hpp"
#include "kernels.hpp"
#include "lbm_hip.h"
#include "solver_impl.hpp"

using namespace lbmhip;

struct lbm_strip {
  int nx = 0, nyl = 0, prec = 0, ghost = 1;
  double tau = 0.8, fx = 0.0, fy = 0.0, u_in = 0.0;
  bool open_x = false;
  int block_x = 512, rows = 1;  // the plain solver's fp32/fp64 defaults (LBM_BLOCK overrides)
  // two-step tiles: 256 columns (float and double), height chosen by step2_tile_rows() with at
  // least two full waves of tiles when rows2 = -1 (16 rows at 8192 x 1024: 2,048 tiles on four
  // 256-thread workgroups per CU); XCD-row tile order (whole tile rows per XCD; round-robin
  // where the tile rows are not a multiple of 8). 8192 x 1024 fp32: 256 x 16 136.1-136.3k
  // MLUPS (XCD rows), 134.5k (round-robin), 131.5k (XCD bands); 512 x 16 132.4-134.2k (XCD
  // rows), 131.3k (round-robin); 512 x 32 (one wave) 133.1k
  int block2_x = 256, rows2 = -1, order2 = 2;
  mutable int rows2_auto = 0;  // the automatic height for the owned rows (0: not yet known)
  std::vector<uint8_t> solid;  // rows() * nx, ghost rows included
  size_t S = 0;                // plane stride (values) of the device buffers
  void* g = nullptr;
  void* gn = nullptr;
  uint16_t* links = nullptr;
  uint8_t* segflags = nullptr;
  int nseg = 0;
  void* halo = nullptr;  // pinned host staging: send up, send down, recv down, recv up
  hipStream_t streams[2] = {nullptr, nullptr};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  int rows_total() const { return nyl + 2 * ghost; }
  size_t cells() const { return static_cast<size_t>(nx) * rows_total(); }
  size_t rs() const { return real_size(prec); }
  size_t bytes() const { return S * kQ * rs(); }
  size_t halo_values() const { return static_cast<size_t>(halo_entries(ghost)) * nx; }
  char* plane_row(void* buf, int i, int y) const {
    return static_cast<char*>(buf) + (i * S + static_cast<size_t>(y) * nx) * rs();
  }

  int create(int nx_, int nyl_, int prec_, int ghost_) {
    nx = nx_;
    nyl = nyl_;
    prec = prec_;
    ghost = ghost_;
    block2_x = 256;
    S = plane_stride(cells(), rs());
    if (S * rs() >= (size_t(1) << 32))
      return fail(LBM_EINVAL, "one population plane of the strip must stay below 4 GiB");
    LBM_TRY(parse_block_env(&block_x, &rows));
    for (auto& s : streams) LBM_HIP_TRY(hipStreamCreate(&s));
    LBM_HIP_TRY(hipEventCreate(&ev0));
    LBM_HIP_TRY(hipEventCreate(&ev1));
    nseg = (nx + 63) / 64;
    // whole dwords: the two-step kernel reads a flag with a scalar load of its aligned dword
    const size_t flag_bytes = (static_cast<size_t>(rows_total()) * nseg + 3) & ~size_t(3);
    if (hipMalloc(&g, bytes()) != hipSuccess || hipMalloc(&gn, bytes()) != hipSuccess ||
        hipMalloc(&links, cells() * sizeof(uint16_t)) != hipSuccess ||
        hipMalloc(&segflags, flag_bytes) != hipSuccess)
      return fail(LBM_ENOMEM, "device allocation of the strip failed");
    LBM_HIP_TRY(hipMemset(g, 0, bytes()));
    LBM_HIP_TRY(hipMemset(gn, 0, bytes()));
    LBM_HIP_TRY(hipMemset(segflags, 0, flag_bytes));
    if (hipHostMalloc(&halo, 4 * halo_values() * rs(), hipHostMallocDefault) != hipSuccess)
      return fail(LBM_ENOMEM, "pinned host allocation of the halo buffers failed");
    solid.assign(cells(), 0);
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

  int sync_all() {
    for (auto s : streams) LBM_HIP_TRY(hipStreamSynchronize(s));
    return LBM_OK;
  }

  int upload_links() {
    if (open_x && nx < 3) return fail(LBM_EINVAL, "an open domain needs nx >= 3");
    // rows 0 and rows_total() - 1 get only the solid bit; every other row its link codes (with
    // two ghost rows, the inner ones are updated by the first of two fused steps)
    const Links L = build_links(nx, rows_total(), solid.data(), open_x, false);
    LBM_TRY(sync_all());
    LBM_HIP_TRY(hipMemcpy(links, L.code.data(), L.code.size() * sizeof(uint16_t),
                          hipMemcpyHostToDevice));
    LBM_HIP_TRY(hipMemcpy(segflags, L.seg.data(), L.seg.size(), hipMemcpyHostToDevice));
    return LBM_OK;
  }

  // Host layout (9 planes of cells() values, contiguous) to / from a device buffer.
  int upload(void* dev, const void* host) {
    const size_t row = cells() * rs();
    for (int i = 0; i < kQ; ++i)
      LBM_HIP_TRY(hipMemcpy(static_cast<char*>(dev) + i * S * rs(),
                            static_cast<const char*>(host) + i * row, row, hipMemcpyHostToDevice));
    return LBM_OK;
  }
  int download(void* host, const void* dev) {
    const size_t row = cells() * rs();
    for (int i = 0; i < kQ; ++i)
      LBM_HIP_TRY(hipMemcpy(static_cast<char*>(host) + i * row,
                            static_cast<const char*>(dev) + i * S * rs(), row,
                            hipMemcpyDeviceToHost));
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
    p.N = S;
    p.nx = nx;
    p.ny = rows_total();
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

  // Parameters of a two-step launch over local rows [first, last] (the walk's tiles).
  template <class T>
  StepParams<T> params2(int first, int last) const {
    StepParams<T> p = params<T>(first, last);
    p.tiles_x = (nx + block2_x - 1) / block2_x;
    if (rows2 > 0) {
      p.rows_per_tile = rows2;
    } else {
      // an occupancy query: computed once for the whole owned range (partial ranges of the
      // overlapped schedule use the same height)
      if (!rows2_auto) {
        StepParams<T> all = params<T>(ghost, ghost + nyl - 1);
        all.tiles_x = p.tiles_x;
        rows2_auto = step2_tile_rows(all, block2_x, 2);
      }
      p.rows_per_tile = rows2_auto;
    }
    const int tiles_y = (p.nrows + p.rows_per_tile - 1) / p.rows_per_tile;
    p.xcd_bands = order2 == 2 && tiles_y % 8 ? 0 : order2;  // XCD rows need groups of 8
    return p;
  }

  bool owned(int first, int last) const { return first >= ghost && last <= ghost + nyl - 1; }

  int compute_rows(int first, int last, int s) {
    if (last < first) return LBM_OK;
    if (!owned(first, last) || s < 0 || s > 1)
      return fail(LBM_EINVAL, "compute_rows: rows must be owned local rows (" +
                                  std::to_string(ghost) + ".." + std::to_string(ghost + nyl - 1) +
                                  "), stream 0 or 1");
    const hipError_t e = prec == LBM_DOUBLE
                             ? launch_step(params<double>(first, last), open_x, block_x, streams[s])
                             : launch_step(params<float>(first, last), open_x, block_x, streams[s]);
    if (e != hipSuccess) return fail(LBM_EHIP, std::string("strip launch: ") + hipGetErrorString(e));
    return LBM_OK;
  }

  int compute_rows2(int first, int last, int s) {
    if (ghost != 2) return fail(LBM_EINVAL, "two steps per launch need a strip with two ghost rows");
    if (open_x) return fail(LBM_EINVAL, "two steps per launch need a periodic x boundary");
    if (last < first) return LBM_OK;
    if (!owned(first, last) || s < 0 || s > 1)
      return fail(LBM_EINVAL, "compute_rows2: rows must be owned local rows (2.." +
                                  std::to_string(nyl + 1) + "), stream 0 or 1");
    const hipError_t e = prec == LBM_DOUBLE
                             ? launch_step2(params2<double>(first, last), block2_x, streams[s])
                             : launch_step2(params2<float>(first, last), block2_x, streams[s]);
    if (e != hipSuccess)
      return fail(LBM_EHIP, std::string("strip two-step launch: ") + hipGetErrorString(e));
    return LBM_OK;
  }

  // The halo targets of a two-step launch: the next-state buffers of the strips above and
  // below (this strip itself when the ring is closed on it).
  template <class T>
  HaloOut<T> halo_to(const lbm_strip& up, const lbm_strip& down) const {
    HaloOut<T> h{};
    h.up = static_cast<T*>(up.gn);
    h.down = static_cast<T*>(down.gn);
    h.up_stride = static_cast<unsigned>(up.S * sizeof(T));
    h.down_stride = static_cast<unsigned>(down.S * sizeof(T));
    h.up_tail = static_cast<unsigned>(up.cells() * sizeof(T));
    h.down_tail = static_cast<unsigned>(down.cells() * sizeof(T));
    h.down_row = down.nyl + 2;
    h.lo = ghost;
    h.hi = ghost + nyl - 1;
    return h;
  }

  int compute_rows2_to(int first, int last, int s, const lbm_strip& up, const lbm_strip& down) {
    if (ghost != 2 || up.ghost != 2 || down.ghost != 2)
      return fail(LBM_EINVAL, "halo delivery needs strips with two ghost rows");
    if (up.nx != nx || down.nx != nx || up.prec != prec || down.prec != prec)
      return fail(LBM_EINVAL, "halo delivery needs strips of the same width and precision");
    if (open_x) return fail(LBM_EINVAL, "two steps per launch need a periodic x boundary");
    if (last < first) return LBM_OK;
    if (!owned(first, last) || s < 0 || s > 1)
      return fail(LBM_EINVAL, "compute_rows2: rows must be owned local rows (2.." +
                                  std::to_string(nyl + 1) + "), stream 0 or 1");
    const hipError_t e =
        prec == LBM_DOUBLE
            ? launch_step2_halo(params2<double>(first, last), halo_to<double>(up, down), block2_x,
                                streams[s])
            : launch_step2_halo(params2<float>(first, last), halo_to<float>(up, down), block2_x,
                                streams[s]);
    if (e != hipSuccess)
      return fail(LBM_EHIP, std::string("strip two-step launch: ") + hipGetErrorString(e));
    return LBM_OK;
  }

  hipError_t ring(hipStream_t st) {
    return prec == LBM_DOUBLE
               ? launch_halo_ring(static_cast<double*>(gn), S, nx, nyl, ghost, st)
               : launch_halo_ring(static_cast<float*>(gn), S, nx, nyl, ghost, st);
  }
};

extern "C" {

int lbm_strip_create_ghost(int nx, int ny_local, int precision, int ghost_rows, lbm_strip** out) {
  if (!out) return fail(LBM_EINVAL, "out handle pointer is null");
  *out = nullptr;
  if (ghost_rows != 1 && ghost_rows != 2) return fail(LBM_EINVAL, "ghost_rows must be 1 or 2");
  if (nx < 1 || ny_local < ghost_rows)
    return fail(LBM_EINVAL, "strip too small (ny_local must be at least ghost_rows)");
  if (precision != LBM_FLOAT && precision != LBM_DOUBLE)
    return fail(LBM_EINVAL, "the strip engine runs in LBM_FLOAT or LBM_DOUBLE");
  auto* s = new lbm_strip();
  const int rc = s->create(nx, ny_local, precision, ghost_rows);
  if (rc != LBM_OK) {
    s->release();
    delete s;
    return rc;
  }
  *out = s;
  return LBM_OK;
}

int lbm_strip_create(int nx, int ny_local, int precision, lbm_strip** out) {
  return lbm_strip_create_ghost(nx, ny_local, precision, 1, out);
}

int lbm_strip_destroy(lbm_strip* s) {
  if (!s) return LBM_OK;
  s->release();
  delete s;
  return LBM_OK;
}

int lbm_strip_halo_values(const lbm_strip* s, size_t* values) {
  if (!s || !values) return fail(LBM_EINVAL, "null strip handle or output");
  *values = s->halo_values();
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
  s->solid.assign(solid_strip, solid_strip + s->cells());
  return s->upload_links();
}

int lbm_strip_upload_state(lbm_strip* s, const void* g) {
  if (!s || !g) return fail(LBM_EINVAL, "null strip handle or state");
  LBM_TRY(s->sync_all());
  LBM_TRY(s->upload(s->g, g));
  return s->upload(s->gn, g);
}

int lbm_strip_download_state(lbm_strip* s, void* g) {
  if (!s || !g) return fail(LBM_EINVAL, "null strip handle or state");
  LBM_TRY(s->sync_all());
  return s->download(g, s->g);
}

int lbm_strip_compute_rows(lbm_strip* s, int first, int last, int stream) {
  if (!s) return fail(LBM_EINVAL, "null strip handle");
  return s->compute_rows(first, last, stream);
}

int lbm_strip_compute_rows2(lbm_strip* s, int first, int last, int stream) {
  if (!s) return fail(LBM_EINVAL, "null strip handle");
  return s->compute_rows2(first, last, stream);
}

int lbm_strip_compute_rows2_to(lbm_strip* s, int first, int last, int stream, lbm_strip* up,
                               lbm_strip* down) {
  if (!s || !up || !down) return fail(LBM_EINVAL, "null strip handle");
  return s->compute_rows2_to(first, last, stream, *up, *down);
}

int lbm_strip_set_two_step_tile(lbm_strip* s, int block_x, int rows, int tile_order) {
  if (!s) return fail(LBM_EINVAL, "null strip handle");
  const int max_x = s->prec == LBM_DOUBLE ? 256 : 512;
  if (block_x != 0 && (block_x < 64 || block_x > max_x || block_x % 64))
    return fail(LBM_EINVAL, "block_x must be 0 or a multiple of 64 up to " + std::to_string(max_x));
  if (rows < -1) return fail(LBM_EINVAL, "rows must be -1 (automatic), 0 (keep) or positive");
  if (tile_order < -1 || tile_order > 2) return fail(LBM_EINVAL, "tile_order must be -1..2");
  if (block_x && block_x != s->block2_x) {
    s->block2_x = block_x;
    s->rows2_auto = 0;
  }
  if (rows) s->rows2 = rows;
  if (tile_order >= 0) s->order2 = tile_order;
  return LBM_OK;
}

int lbm_strip_pack_halo(lbm_strip* s, void* send_up, void* send_down) {
  if (!s || !send_up || !send_down) return fail(LBM_EINVAL, "null strip handle or buffer");
  const size_t row = static_cast<size_t>(s->nx) * s->rs();
  const int n = halo_entries(s->ghost);
  char* up = static_cast<char*>(s->halo);
  char* down = up + n * row;
  hipStream_t st = s->streams[0];  // same stream as the boundary rows
  for (int k = 0; k < n; ++k) {
    int plane = 0, y = 0;
    halo_send(s->ghost, s->nyl, 1, k, plane, y);
    LBM_HIP_TRY(hipMemcpyAsync(up + k * row, s->plane_row(s->gn, plane, y), row,
                               hipMemcpyDeviceToHost, st));
    halo_send(s->ghost, s->nyl, 0, k, plane, y);
    LBM_HIP_TRY(hipMemcpyAsync(down + k * row, s->plane_row(s->gn, plane, y), row,
                               hipMemcpyDeviceToHost, st));
  }
  LBM_HIP_TRY(hipStreamSynchronize(st));  // the interior stream keeps running
  std::memcpy(send_up, up, n * row);
  std::memcpy(send_down, down, n * row);
  return LBM_OK;
}

int lbm_strip_unpack_halo(lbm_strip* s, const void* recv_from_down, const void* recv_from_up) {
  if (!s || !recv_from_down || !recv_from_up) return fail(LBM_EINVAL, "null strip handle or buffer");
  const size_t row = static_cast<size_t>(s->nx) * s->rs();
  const int n = halo_entries(s->ghost);
  char* from_down = static_cast<char*>(s->halo) + 2 * n * row;
  char* from_up = from_down + n * row;
  hipStream_t st = s->streams[0];
  // the staging may still feed the previous exchange's copies
  LBM_HIP_TRY(hipStreamSynchronize(st));
  std::memcpy(from_down, recv_from_down, n * row);
  std::memcpy(from_up, recv_from_up, n * row);
  for (int k = 0; k < n; ++k) {
    int plane = 0, y = 0;
    halo_recv(s->ghost, s->nyl, 1, k, plane, y);
    LBM_HIP_TRY(hipMemcpyAsync(s->plane_row(s->gn, plane, y), from_down + k * row, row,
                               hipMemcpyHostToDevice, st));
    halo_recv(s->ghost, s->nyl, 0, k, plane, y);
    LBM_HIP_TRY(hipMemcpyAsync(s->plane_row(s->gn, plane, y), from_up + k * row, row,
                               hipMemcpyHostToDevice, st));
  }
  return LBM_OK;
}

int lbm_strip_finish_step(lbm_strip* s) {
  if (!s) return fail(LBM_EINVAL, "null strip handle");
  LBM_TRY(s->sync_all());
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
  // The strip at its real shape with the ring closed on itself, all on stream 0: after each
  // single step the ghost rows of the new state are refreshed from the strip's own owned rows
  // by one device kernel (the contents an exchange with the neighbours would deliver); with two
  // ghost rows and a periodic x boundary, pairs of steps take one two-step launch that stores
  // its halo entries into its own next-state ghost rows (lbm_strip_compute_rows2_to with the
  // strip as both neighbours): no refresh kernel.
  if (!s || steps < 0) return fail(LBM_EINVAL, "null strip handle or negative steps");
  hipStream_t st = s->streams[0];
  LBM_TRY(s->sync_all());
  const int first = s->ghost, last = s->ghost + s->nyl - 1;
  const bool two = s->ghost == 2 && !s->open_x;
  LBM_HIP_TRY(hipEventRecord(s->ev0, st));
  for (int k = 0; k < steps;) {
    const bool pair = two && steps - k >= 2;
    if (pair) {
      LBM_TRY(s->compute_rows2_to(first, last, 0, *s, *s));
    } else {
      const hipError_t e =
          s->prec == LBM_DOUBLE
              ? launch_step(s->params<double>(first, last), s->open_x, s->block_x, st)
              : launch_step(s->params<float>(first, last), s->open_x, s->block_x, st);
      if (e != hipSuccess)
        return fail(LBM_EHIP, std::string("strip launch: ") + hipGetErrorString(e));
      const hipError_t h = s->ring(st);
      if (h != hipSuccess)
        return fail(LBM_EHIP, std::string("halo launch: ") + hipGetErrorString(h));
    }
    std::swap(s->g, s->gn);
    k += pair ? 2 : 1;
  }
  LBM_HIP_TRY(hipEventRecord(s->ev1, st));
  LBM_HIP_TRY(hipEventSynchronize(s->ev1));
  float ms = 0.0f;
  LBM_HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  if (seconds) *seconds = ms * 1e-3;
  return LBM_OK;
}

}  // extern "C"
