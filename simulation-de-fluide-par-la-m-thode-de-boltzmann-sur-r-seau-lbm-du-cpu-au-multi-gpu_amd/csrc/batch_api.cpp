// C ABI of the batched persistent step (lbm_batch_*): host-side layout of the domains, the
// descriptor table and tile work-list, HIP-graph capture of multi-step loops and the autotuner
// of the launch shape. No reference counterpart.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <queue>
#include <string>
#include <vector>

#include "batch.hpp"
#include "d2q9.hpp"
#include "host_common.hpp"
#include "kernels.hpp"
#include "lbm_hip.h"

using namespace lbmhip;

namespace {

struct DomainHost {
  int nx, ny;
  double tau, fx, fy, rho0, ux0, uy0;
  std::vector<uint8_t> solid;
  size_t offset = 0, cell_offset = 0, seg_offset = 0;
  int nseg = 0;
  size_t N() const { return static_cast<size_t>(nx) * ny; }
  // plane stride: the cells rounded up to 64, plus the two-step kernel's discard slots
  size_t S() const { return (N() + 63) / 64 * 64 + kPlaneTail; }
};

constexpr size_t kAlignElems = 64;  // plane-0 offsets aligned to 64 elements

}  // namespace

struct lbm_batch {
  int prec = LBM_FLOAT;
  std::vector<DomainHost> doms;
  size_t total_elems = 0, total_cells = 0, total_segs = 0;
  void* g = nullptr;
  void* gn = nullptr;
  uint16_t* links = nullptr;
  uint8_t* segflags = nullptr;
  void* ddesc = nullptr;
  BatchTile* dtiles = nullptr;
  size_t tiles_cap = 0;
  BatchSeg* dsegs = nullptr;  // wave work-list (order 2)
  size_t segs_cap = 0;
  int ntiles = 0;             // entries of the list the current order walks
  int tile_x = 256, tile_rows = 4, workgroups = 0;
  // Two steps per launch: per-domain sub-tiles of at most tile2_x columns and band height
  // rows2 (tile2_rows, or the modelled choice when it is 0), packed into work items of block2
  // threads (build_items2).
  int steps_per_launch = 2;
  int block2 = 512;  // threads per workgroup of the two-step kernel (double: 256)
  int tile2_x = 512, tile2_rows = 0, rows2 = 0;
  int workgroups2 = 0;  // persistent grid of the two-step kernel (0: one workgroup per item)
  BatchWave* dwaves = nullptr;
  size_t waves_cap = 0;
  int nitems2 = 0;
  double fill2 = 0.0;   // cell rows / lane rows of the items
  int order = 0, store_nt = 1;  // tile order (see batch.hip) and non-temporal stores
  bool one_rsrc = true;         // every domain's nine planes span < 4 GiB
  int graph_steps = 100;
  int parity = 0;  // 0: current state in g
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void* staging = nullptr;
  size_t staging_bytes = 0;

  size_t rs() const { return real_size(prec); }

  void drop_graphs() {
    for (auto& e : exec)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
  }

  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    for (void* p : {g, gn, static_cast<void*>(links), static_cast<void*>(segflags), ddesc,
                    static_cast<void*>(dtiles), static_cast<void*>(dsegs),
                    static_cast<void*>(dwaves), staging})
      if (p) (void)hipFree(p);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  template <class T>
  int upload_descriptors() {
    std::vector<BatchDomain<T>> d(doms.size());
    for (size_t k = 0; k < doms.size(); ++k) {
      const DomainHost& h = doms[k];
      BatchDomain<T>& b = d[k];
      b.offset = h.offset;
      b.cell_offset = h.cell_offset;
      b.seg_offset = h.seg_offset;
      b.N = h.S();
      b.nx = h.nx;
      b.ny = h.ny;
      b.nseg = h.nseg;
      b.pad = 0;
      const T tau = static_cast<T>(h.tau);
      b.omega = T(1) / tau;  // as the single-domain step (solver_v3.cu:252-253)
      b.fp = T(1) - T(0.5) * b.omega;
      b.fx = static_cast<T>(h.fx);
      b.fy = static_cast<T>(h.fy);
    }
    LBM_HIP_TRY(hipMemcpy(ddesc, d.data(), d.size() * sizeof(BatchDomain<T>), hipMemcpyHostToDevice));
    return LBM_OK;
  }

  // Order 2: per-wavefront segments of 64 columns over tile_rows rows, domain by domain, row
  // block by row block (neighbouring segments go to neighbouring waves of one workgroup).
  int build_segs() {
    std::vector<BatchSeg> t;
    for (size_t k = 0; k < doms.size(); ++k)
      for (int y = 0; y < doms[k].ny; y += tile_rows)
        for (int sg = 0; sg < doms[k].nseg; ++sg)
          t.push_back(BatchSeg{static_cast<int>(k), sg, y, std::min(y + tile_rows, doms[k].ny)});
    ntiles = static_cast<int>(t.size());
    if (t.size() > segs_cap) {
      if (dsegs) LBM_HIP_TRY(hipFree(dsegs));
      dsegs = nullptr;
      if (hipMalloc(&dsegs, t.size() * sizeof(BatchSeg)) != hipSuccess)
        return fail(LBM_ENOMEM, "device allocation of the segment list failed");
      segs_cap = t.size();
    }
    LBM_HIP_TRY(hipMemcpy(dsegs, t.data(), t.size() * sizeof(BatchSeg), hipMemcpyHostToDevice));
    drop_graphs();
    return LBM_OK;
  }

  int build_tiles() {
    if (order == 2) return build_segs();
    std::vector<BatchTile> t;
    for (size_t k = 0; k < doms.size(); ++k) {
      const int tx = (doms[k].nx + tile_x - 1) / tile_x, ty = (doms[k].ny + tile_rows - 1) / tile_rows;
      for (int y = 0; y < ty; ++y)
        for (int x = 0; x < tx; ++x) t.push_back(BatchTile{static_cast<int>(k), x, y, 0});
    }
    ntiles = static_cast<int>(t.size());
    if (t.size() > tiles_cap) {
      if (dtiles) LBM_HIP_TRY(hipFree(dtiles));
      dtiles = nullptr;
      if (hipMalloc(&dtiles, t.size() * sizeof(BatchTile)) != hipSuccess)
        return fail(LBM_ENOMEM, "device allocation of the tile list failed");
      tiles_cap = t.size();
    }
    LBM_HIP_TRY(hipMemcpy(dtiles, t.data(), t.size() * sizeof(BatchTile), hipMemcpyHostToDevice));
    drop_graphs();
    return LBM_OK;
  }

  // ---- two-step work items ----
  struct Sub {
    int domain, x0, y0, rows, width;
  };
  struct Item {
    int nr_max, free;
    std::vector<int> subs;
  };

  // Sub-tiles: domain k's columns in tiles of tile2_x columns and one last tile of the remainder
  // (its width rounded up to 64: the remainders of different domains pack together), its rows
  // in ceil(ny / Hb) bands of near-equal height.
  std::vector<Sub> subtiles(int Hb) const {
    std::vector<Sub> s;
    for (size_t k = 0; k < doms.size(); ++k) {
      const int nx = doms[k].nx, ny = doms[k].ny;
      const int nty = (ny + Hb - 1) / Hb, H = (ny + nty - 1) / nty;
      for (int y = 0; y < ny; y += H)
        for (int x = 0; x < nx; x += tile2_x)
          s.push_back(Sub{static_cast<int>(k), x, y, std::min(H, ny - y),
                          std::min(tile2_x, (nx - x + 63) / 64 * 64)});
    }
    return s;
  }

  // First-fit packing, tallest sub-tiles first, into items of block2 lanes: a sub-tile joins the
  // first open item with room whose tallest sub-tile is at most kSlack rows taller (the item
  // lasts as long as its tallest walk), else opens one. The items come out tallest first, the
  // order the workgroups are dispatched in: the short ones fill the last round.
  static constexpr int kSlack = 4;
  std::vector<Item> pack(std::vector<Sub>& s) const {
    std::stable_sort(s.begin(), s.end(), [](const Sub& a, const Sub& b) {
      return a.rows != b.rows ? a.rows > b.rows : a.width > b.width;
    });
    std::vector<Item> items;
    std::vector<size_t> open;
    for (size_t i = 0; i < s.size(); ++i) {
      const int nr = s[i].rows + 2;
      open.erase(std::remove_if(open.begin(), open.end(),
                                [&](size_t j) { return items[j].nr_max - nr > kSlack || items[j].free < 64; }),
                 open.end());
      bool placed = false;
      for (size_t j : open)
        if (items[j].free >= s[i].width) {
          items[j].free -= s[i].width;
          items[j].subs.push_back(static_cast<int>(i));
          placed = true;
          break;
        }
      if (!placed) {
        items.push_back(Item{nr, block2 - s[i].width, {static_cast<int>(i)}});
        open.push_back(items.size() - 1);
      }
    }
    return items;
  }

  // Modelled time of the items on `slots` resident workgroups, dispatched in order each to the
  // first free slot: an item lasts its tallest walk's step-1 rows.
  static double makespan(const std::vector<Item>& items, long slots) {
    std::priority_queue<double, std::vector<double>, std::greater<double>> free_at;
    for (long k = 0; k < slots; ++k) free_at.push(0.0);
    double end = 0.0;
    for (const Item& it : items) {
      const double t0 = free_at.top();
      free_at.pop();
      free_at.push(t0 + it.nr_max);
      end = std::max(end, t0 + it.nr_max);
    }
    return end;
  }

  // Band height when tile2_rows is 0: the candidate 16 .. 96 whose packed items the schedule
  // model finishes soonest (ties: the taller, fewer rows re-read), on the resident workgroups
  // of the two-step kernel.
  int model_rows() const {
    int per_cu = prec == LBM_DOUBLE ? batch_step2_resident<double>(block2, store_nt != 0, one_rsrc)
                                    : batch_step2_resident<float>(block2, store_nt != 0, one_rsrc);
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const long slots = static_cast<long>(std::max(per_cu, 1)) * std::max(cus, 1);
    int best = 32;
    double best_t = 0.0;
    for (int Hb = 96; Hb >= 16; Hb -= 4) {
      std::vector<Sub> s = subtiles(Hb);
      const double tm = makespan(pack(s), slots);
      if (best_t == 0.0 || tm < best_t * 0.999) {
        best_t = tm;
        best = Hb;
      }
    }
    return best;
  }

  // The two-step list: one entry per wave of every item (BatchWave).
  int build_tiles2() {
    rows2 = tile2_rows > 0 ? tile2_rows : model_rows();
    std::vector<Sub> s = subtiles(rows2);
    const std::vector<Item> items = pack(s);
    const int nw = block2 / 64;
    const size_t lds_cap = batch_step2_lds_bytes(block2, rs()) / rs();
    std::vector<BatchWave> w(items.size() * nw);
    double cell_rows = 0.0, lane_rows = 0.0;
    for (size_t i = 0; i < items.size(); ++i) {
      BatchWave* e = w.data() + i * nw;
      int lane = 0;
      size_t lds = 0;
      for (int j : items[i].subs) {
        const Sub& u = s[j];
        for (int v = lane / 64; v < (lane + u.width) / 64; ++v)
          e[v] = BatchWave{u.domain, u.x0, u.y0, u.rows, lane, u.width, static_cast<int>(lds), items[i].nr_max};
        lane += u.width;
        lds += step2_region_elems(u.width);
        cell_rows += static_cast<double>(std::min(u.width, doms[u.domain].nx - u.x0)) * (u.rows + 2);
      }
      if (lane > block2 || lds > lds_cap) return fail(LBM_EINVAL, "two-step item packing overflow");
      for (int v = lane / 64; v < nw; ++v) e[v] = BatchWave{-1, 0, 0, 0, 0, 0, 0, items[i].nr_max};
      lane_rows += static_cast<double>(block2) * items[i].nr_max;
    }
    fill2 = lane_rows > 0.0 ? cell_rows / lane_rows : 0.0;
    nitems2 = static_cast<int>(items.size());
    if (w.size() > waves_cap) {
      if (dwaves) LBM_HIP_TRY(hipFree(dwaves));
      dwaves = nullptr;
      if (hipMalloc(&dwaves, w.size() * sizeof(BatchWave)) != hipSuccess)
        return fail(LBM_ENOMEM, "device allocation of the two-step work list failed");
      waves_cap = w.size();
    }
    LBM_HIP_TRY(hipMemcpy(dwaves, w.data(), w.size() * sizeof(BatchWave), hipMemcpyHostToDevice));
    drop_graphs();
    return LBM_OK;
  }

  bool two_step() const { return steps_per_launch == 2; }

  template <class T>
  BatchParams<T> params2(bool flip) const {
    BatchParams<T> p = params<T>(flip);
    p.waves = dwaves;
    p.ntiles = nitems2;
    return p;
  }

  int launch2(bool flip) {
    const hipError_t e = prec == LBM_DOUBLE
                             ? launch_batch_step2(params2<double>(flip), block2, workgroups2, stream)
                             : launch_batch_step2(params2<float>(flip), block2, workgroups2, stream);
    if (e != hipSuccess) return fail(LBM_EHIP, std::string("two-step batch launch: ") + hipGetErrorString(e));
    return LBM_OK;
  }

  // Launches advancing `steps` steps from buffer parity par (updated): pairs in two-step
  // launches when on, one-step launches otherwise and for an odd remainder.
  int issue(int steps, int& par) {
    if (two_step())
      for (; steps >= 2; steps -= 2) {
        LBM_TRY(launch2(par == 1));
        par ^= 1;
      }
    for (; steps > 0; --steps) {
      LBM_TRY(launch(par == 1));
      par ^= 1;
    }
    return LBM_OK;
  }

  // Parity change of a graph of graph_steps steps (one per launch).
  int graph_flip() const {
    const int launches = two_step() ? graph_steps / 2 + (graph_steps & 1) : graph_steps;
    return launches & 1;
  }

  template <class T>
  BatchParams<T> params(bool flip) const {
    BatchParams<T> p{};
    p.g = static_cast<const T*>(flip ? gn : g);
    p.gn = static_cast<T*>(flip ? g : gn);
    p.links = links;
    p.segflags = segflags;
    p.domains = static_cast<const BatchDomain<T>*>(ddesc);
    p.tiles = dtiles;
    p.segs = dsegs;
    p.ndomains = static_cast<int>(doms.size());
    p.ntiles = ntiles;
    p.tile_rows = tile_rows;
    const int wg = std::max(1, std::min(workgroups, order == 2 ? (ntiles + tile_x / 64 - 1) / (tile_x / 64) : ntiles));
    p.chunk = (ntiles + wg - 1) / wg;
    p.order = order;
    p.store_nt = store_nt;
    p.one_rsrc = one_rsrc ? 1 : 0;
    return p;
  }

  int launch(bool flip) {
    // order 2 needs no more workgroups than it has waves' worth of segments
    const int wg = std::max(1, std::min(workgroups, order == 2 ? (ntiles + tile_x / 64 - 1) / (tile_x / 64) : ntiles));
    const hipError_t e = prec == LBM_DOUBLE
                             ? launch_batch(params<double>(flip), tile_x, wg, stream)
                             : launch_batch(params<float>(flip), tile_x, wg, stream);
    if (e != hipSuccess) return fail(LBM_EHIP, std::string("batch launch: ") + hipGetErrorString(e));
    return LBM_OK;
  }

  int step_direct() {
    LBM_TRY(launch(parity == 1));
    parity ^= 1;
    return LBM_OK;
  }

  // A graph of graph_steps steps starting from parity p; launches are captured, not run.
  int build_graph(int p) {
    hipGraph_t graph = nullptr;
    LBM_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int par = p;
    const int rc = issue(graph_steps, par);
    const hipError_t e = hipStreamEndCapture(stream, &graph);
    if (rc != LBM_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return fail(LBM_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(&exec[p], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) return fail(LBM_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e2));
    return LBM_OK;
  }

  int run_async(int steps) {
    while (steps >= graph_steps && graph_steps > 1) {
      if (!exec[parity]) LBM_TRY(build_graph(parity));
      LBM_HIP_TRY(hipGraphLaunch(exec[parity], stream));
      parity ^= graph_flip();
      steps -= graph_steps;
    }
    return issue(steps, parity);
  }

  void* current() const { return parity == 0 ? g : gn; }
};

extern "C" {

int lbm_batch_create(int n, const lbm_domain_desc* domains, int precision, lbm_batch** out) {
  if (!out) return fail(LBM_EINVAL, "out handle pointer is null");
  *out = nullptr;
  if (n < 1 || !domains) return fail(LBM_EINVAL, "need at least one domain");
  if (precision != LBM_FLOAT && precision != LBM_DOUBLE)
    return fail(LBM_EINVAL, "the batched step runs in LBM_FLOAT or LBM_DOUBLE");
  auto* b = new lbm_batch();
  b->prec = precision;
  auto bail = [&](int rc) {
    b->release();
    delete b;
    return rc;
  };
  b->doms.resize(n);
  for (int k = 0; k < n; ++k) {
    const lbm_domain_desc& d = domains[k];
    if (d.nx < 1 || d.ny < 1 || !d.solid) return bail(fail(LBM_EINVAL, "domain " + std::to_string(k) + ": invalid grid or mask"));
    const bool ok = precision == LBM_DOUBLE ? d.tau > 0.5 : static_cast<float>(d.tau) > 0.5f;
    if (!ok) return bail(fail(LBM_EINVAL, "domain " + std::to_string(k) + ": tau must be > 0.5"));
    DomainHost& h = b->doms[k];
    h.nx = d.nx;
    h.ny = d.ny;
    h.tau = d.tau;
    h.fx = d.force_x;
    h.fy = d.force_y;
    h.rho0 = d.rho0;
    h.ux0 = d.ux0;
    h.uy0 = d.uy0;
    h.solid.assign(d.solid, d.solid + h.N());
    h.nseg = (d.nx + 63) / 64;
    h.offset = b->total_elems;
    h.cell_offset = b->total_cells;
    h.seg_offset = b->total_segs;
    b->total_elems += (kQ * h.S() + kAlignElems - 1) / kAlignElems * kAlignElems;
    b->total_cells += h.N();
    b->total_segs += static_cast<size_t>(h.ny) * h.nseg;
    if (kQ * h.S() * b->rs() >= (size_t(1) << 32)) b->one_rsrc = false;
    if (h.S() * b->rs() >= (size_t(1) << 32))
      return bail(fail(LBM_EINVAL, "domain " + std::to_string(k) + ": a plane must stay below 4 GiB"));
  }
  int rc = parse_block_env(&b->tile_x, &b->tile_rows);
  if (rc != LBM_OK) return bail(rc);
  if (hipStreamCreate(&b->stream) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
      hipEventCreate(&b->ev1) != hipSuccess)
    return bail(fail(LBM_EHIP, "stream / event creation failed"));
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  b->workgroups = std::max(1, cus) * 8;
  // the two-step kernel's LDS ring holds block2 + 2 columns of six planes per ring row
  b->block2 = b->tile2_x = precision == LBM_DOUBLE ? 256 : 512;
  const size_t bytes = b->total_elems * b->rs();
  // segment flags in whole dwords: the two-step kernel reads a flag with a scalar dword load
  if (hipMalloc(&b->g, bytes) != hipSuccess || hipMalloc(&b->gn, bytes) != hipSuccess ||
      hipMalloc(&b->links, b->total_cells * sizeof(uint16_t)) != hipSuccess ||
      hipMalloc(&b->segflags, (b->total_segs + 3) & ~size_t(3)) != hipSuccess ||
      hipMalloc(&b->ddesc, n * std::max(sizeof(BatchDomain<double>), sizeof(BatchDomain<float>))) != hipSuccess)
    return bail(fail(LBM_ENOMEM, "device allocation of the batch failed"));
  // link codes, flags and the initial state of every domain
  std::vector<uint16_t> codes(b->total_cells);
  std::vector<uint8_t> segs(b->total_segs);
  std::vector<char> state(bytes, 0);
  for (auto& h : b->doms) {
    const Links L = build_links(h.nx, h.ny, h.solid.data(), false, true);
    std::copy(L.code.begin(), L.code.end(), codes.begin() + h.cell_offset);
    std::copy(L.seg.begin(), L.seg.end(), segs.begin() + h.seg_offset);
    // the preimage is computed with planes nx * ny apart, then laid out at the plane stride
    auto place = [&](auto* R, auto* dst) {
      for (int i = 0; i < kQ; ++i) std::copy(R + i * h.N(), R + (i + 1) * h.N(), dst + i * h.S());
    };
    if (precision == LBM_DOUBLE) {
      double eq[kQ];
      equilibrium_state<double>(h.rho0, h.ux0, h.uy0, eq);
      std::vector<double> R(kQ * h.N());
      preimage_uniform(h.nx, h.ny, h.solid.data(), false, eq, R.data());
      place(R.data(), reinterpret_cast<double*>(state.data()) + h.offset);
    } else {
      float eq[kQ];
      equilibrium_state<float>(static_cast<float>(h.rho0), static_cast<float>(h.ux0),
                               static_cast<float>(h.uy0), eq);
      std::vector<float> R(kQ * h.N());
      preimage_uniform(h.nx, h.ny, h.solid.data(), false, eq, R.data());
      place(R.data(), reinterpret_cast<float*>(state.data()) + h.offset);
    }
  }
  if (hipMemcpy(b->links, codes.data(), codes.size() * sizeof(uint16_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->segflags, segs.data(), segs.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->g, state.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->gn, state.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(LBM_EHIP, "upload of the batch failed"));
  rc = precision == LBM_DOUBLE ? b->upload_descriptors<double>() : b->upload_descriptors<float>();
  if (rc == LBM_OK) rc = b->build_tiles();
  if (rc == LBM_OK) rc = b->build_tiles2();
  if (rc != LBM_OK) return bail(rc);
  *out = b;
  return LBM_OK;
}

int lbm_batch_destroy(lbm_batch* b) {
  if (!b) return LBM_OK;
  b->release();
  delete b;
  return LBM_OK;
}

int lbm_batch_set_graph_steps(lbm_batch* b, int steps_per_graph) {
  if (!b || steps_per_graph < 1) return fail(LBM_EINVAL, "need a batch and steps_per_graph >= 1");
  LBM_HIP_TRY(hipStreamSynchronize(b->stream));
  b->graph_steps = steps_per_graph;
  b->drop_graphs();
  return LBM_OK;
}

int lbm_batch_set_shape(lbm_batch* b, int tile_x, int tile_rows, int workgroups) {
  if (!b) return fail(LBM_EINVAL, "null batch handle");
  LBM_TRY(check_block(tile_x, tile_rows));
  if (workgroups < 1) return fail(LBM_EINVAL, "workgroups must be >= 1");
  LBM_HIP_TRY(hipStreamSynchronize(b->stream));
  b->tile_x = tile_x;
  b->tile_rows = tile_rows;
  b->workgroups = workgroups;
  return b->build_tiles();
}

int lbm_batch_set_order(lbm_batch* b, int order, int streaming_stores) {
  if (!b) return fail(LBM_EINVAL, "null batch handle");
  if (order < 0 || order > 2) return fail(LBM_EINVAL, "order must be 0, 1 or 2");
  LBM_HIP_TRY(hipStreamSynchronize(b->stream));
  b->order = order;
  b->store_nt = streaming_stores ? 1 : 0;
  return b->build_tiles();
}

int lbm_batch_get_order(const lbm_batch* b, int* order, int* streaming_stores) {
  if (!b) return fail(LBM_EINVAL, "null batch handle");
  if (order) *order = b->order;
  if (streaming_stores) *streaming_stores = b->store_nt;
  return LBM_OK;
}

int lbm_batch_set_steps_per_launch(lbm_batch* b, int steps_per_launch, int tile_x, int tile_rows,
                                   int workgroups) {
  if (!b) return fail(LBM_EINVAL, "null batch handle");
  if (steps_per_launch != 1 && steps_per_launch != 2)
    return fail(LBM_EINVAL, "steps_per_launch must be 1 or 2");
  if (tile_x != 0 && (tile_x < 64 || tile_x > b->block2 || tile_x % 64))
    return fail(LBM_EINVAL, "tile_x must be 0 or a multiple of 64 up to " + std::to_string(b->block2) +
                                " for this precision");
  if (tile_rows < -1) return fail(LBM_EINVAL, "tile_rows must be -1 (modelled), 0 (keep) or positive");
  if (workgroups < -1) return fail(LBM_EINVAL, "workgroups must be -1 (keep), 0 or positive");
  LBM_HIP_TRY(hipStreamSynchronize(b->stream));
  b->steps_per_launch = steps_per_launch;
  if (tile_x) b->tile2_x = tile_x;
  if (tile_rows) b->tile2_rows = tile_rows < 0 ? 0 : tile_rows;
  if (workgroups >= 0) b->workgroups2 = workgroups;
  b->drop_graphs();
  return b->build_tiles2();
}

int lbm_batch_get_steps_per_launch(const lbm_batch* b, int* steps_per_launch, int* tile_x,
                                   int* tile_rows, int* workgroups) {
  if (!b) return fail(LBM_EINVAL, "null batch handle");
  if (steps_per_launch) *steps_per_launch = b->two_step() ? 2 : 1;
  if (tile_x) *tile_x = b->tile2_x;
  if (tile_rows) *tile_rows = b->rows2;
  if (workgroups) *workgroups = b->workgroups2;
  return LBM_OK;
}

int lbm_batch_set_two_step_threads(lbm_batch* b, int threads) {
  if (!b) return fail(LBM_EINVAL, "null batch handle");
  const int most = b->prec == LBM_DOUBLE ? 256 : 512;
  if (threads < 64 || threads > most || threads % 64)
    return fail(LBM_EINVAL, "threads must be a multiple of 64 up to " + std::to_string(most) +
                                " for this precision");
  LBM_HIP_TRY(hipStreamSynchronize(b->stream));
  b->block2 = threads;
  b->tile2_x = std::min(b->tile2_x, threads);
  b->drop_graphs();
  return b->build_tiles2();
}

int lbm_batch_two_step_layout(const lbm_batch* b, int* threads, int* items, int* band_rows,
                              int* modelled, double* fill) {
  if (!b) return fail(LBM_EINVAL, "null batch handle");
  if (threads) *threads = b->block2;
  if (items) *items = b->nitems2;
  if (band_rows) *band_rows = b->rows2;
  if (modelled) *modelled = b->tile2_rows == 0 ? 1 : 0;
  if (fill) *fill = b->fill2;
  return LBM_OK;
}

int lbm_batch_run(lbm_batch* b, int steps) {
  if (!b || steps < 0) return fail(LBM_EINVAL, "null batch handle or negative steps");
  LBM_TRY(b->run_async(steps));
  LBM_HIP_TRY(hipStreamSynchronize(b->stream));
  return LBM_OK;
}

int lbm_batch_timed_run(lbm_batch* b, int steps, double* seconds) {
  if (!b || steps < 0) return fail(LBM_EINVAL, "null batch handle or negative steps");
  // Graphs are built before the timed region (instantiation is not part of a step).
  if (steps >= b->graph_steps && b->graph_steps > 1) {
    if (!b->exec[b->parity]) LBM_TRY(b->build_graph(b->parity));
    if (b->graph_flip() && !b->exec[b->parity ^ 1]) LBM_TRY(b->build_graph(b->parity ^ 1));
  }
  LBM_HIP_TRY(hipEventRecord(b->ev0, b->stream));
  LBM_TRY(b->run_async(steps));
  LBM_HIP_TRY(hipEventRecord(b->ev1, b->stream));
  LBM_HIP_TRY(hipEventSynchronize(b->ev1));
  float ms = 0.0f;
  LBM_HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  if (seconds) *seconds = ms * 1e-3;
  return LBM_OK;
}

int lbm_batch_populations(lbm_batch* b, int domain, void* out) {
  if (!b || !out || domain < 0 || domain >= static_cast<int>(b->doms.size()))
    return fail(LBM_EINVAL, "invalid batch handle, domain or output");
  const DomainHost& h = b->doms[domain];
  const size_t bytes = kQ * h.N() * b->rs();
  if (b->staging_bytes < bytes) {
    if (b->staging) LBM_HIP_TRY(hipFree(b->staging));
    b->staging = nullptr;
    b->staging_bytes = 0;
    if (hipMalloc(&b->staging, bytes) != hipSuccess) return fail(LBM_ENOMEM, "staging allocation failed");
    b->staging_bytes = bytes;
  }
  hipError_t e;
  if (b->prec == LBM_DOUBLE) {
    GatherParams<double> p{};
    p.g = static_cast<const double*>(b->current()) + h.offset;
    p.links = b->links + h.cell_offset;
    p.N = h.S();
    p.nx = h.nx;
    p.ny = h.ny;
    p.fx = h.fx;
    p.fy = h.fy;
    e = launch_gather(p, static_cast<double*>(b->staging), b->stream);
  } else {
    GatherParams<float> p{};
    p.g = static_cast<const float*>(b->current()) + h.offset;
    p.links = b->links + h.cell_offset;
    p.N = h.S();
    p.nx = h.nx;
    p.ny = h.ny;
    p.fx = static_cast<float>(h.fx);
    p.fy = static_cast<float>(h.fy);
    e = launch_gather(p, static_cast<float*>(b->staging), b->stream);
  }
  if (e != hipSuccess) return fail(LBM_EHIP, std::string("gather launch: ") + hipGetErrorString(e));
  LBM_HIP_TRY(hipStreamSynchronize(b->stream));
  LBM_HIP_TRY(hipMemcpy(out, b->staging, bytes, hipMemcpyDeviceToHost));
  return LBM_OK;
}

int lbm_batch_total_cells(const lbm_batch* b, long long* cells) {
  if (!b || !cells) return fail(LBM_EINVAL, "null batch handle or output");
  *cells = static_cast<long long>(b->total_cells);
  return LBM_OK;
}

namespace {

// The timed trials of lbm_batch_autotune (the state is saved and restored around them).
int batch_autotune_trials(lbm_batch* b, int steps, int* tile_x, int* tile_rows, int* workgroups,
                          double* best_mlups) {
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  cus = std::max(cus, 1);
  double best = -1.0;
  int btx = b->tile_x, brows = b->tile_rows, bwg = b->workgroups, border = b->order,
      bnt = b->store_nt;
  for (int ord : {0, 1, 2})
  for (int nt : {1, 0})
  for (int tx : {128, 256, 512})
    for (int r : {1, 2, 4, 8})
      for (int wgm : {4, 8, 16, 32, 64}) {
        b->tile_x = tx;
        b->tile_rows = r;
        b->workgroups = cus * wgm;
        b->order = ord;
        b->store_nt = nt;
        LBM_TRY(b->build_tiles());
        LBM_TRY(b->step_direct());  // warm-up
        LBM_HIP_TRY(hipEventRecord(b->ev0, b->stream));
        for (int k = 0; k < steps; ++k) LBM_TRY(b->step_direct());
        LBM_HIP_TRY(hipEventRecord(b->ev1, b->stream));
        LBM_HIP_TRY(hipEventSynchronize(b->ev1));
        float ms = 0.0f;
        LBM_HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        const double mlups = static_cast<double>(b->total_cells) * steps / (ms * 1e-3) / 1e6;
        if (mlups > best) {
          best = mlups;
          btx = tx;
          brows = r;
          bwg = cus * wgm;
          border = ord;
          bnt = nt;
        }
      }
  b->tile_x = btx;
  b->tile_rows = brows;
  b->workgroups = bwg;
  b->order = border;
  b->store_nt = bnt;
  LBM_TRY(b->build_tiles());
  // The two-step tiles, where run() takes them; best_mlups is then the rate of that path.
  if (b->steps_per_launch == 2) {
    double best2 = -1.0;
    int b2x = b->tile2_x, b2r = b->tile2_rows, b2w = b->workgroups2, b2nt = b->store_nt;
    const int keep_nt = b->store_nt;
    // Sub-tile widths and band heights (0: the modelled height), one workgroup per item; at
    // least 8 launches per candidate, the differences between the best shapes being a few
    // percent.
    const int pairs = std::max(8, (steps + 1) / 2);
    for (int nt : {0, 1})
    for (int tx : {b->block2 / 2, b->block2})
      for (int r : {0, 32, 48, 64}) {
        b->tile2_x = tx;
        b->tile2_rows = r;
        b->workgroups2 = 0;
        b->store_nt = nt;
        LBM_TRY(b->build_tiles2());
        int par = b->parity;
        LBM_TRY(b->issue(2, par));  // warm-up
        LBM_HIP_TRY(hipEventRecord(b->ev0, b->stream));
        LBM_TRY(b->issue(2 * pairs, par));
        LBM_HIP_TRY(hipEventRecord(b->ev1, b->stream));
        LBM_HIP_TRY(hipEventSynchronize(b->ev1));
        b->parity = par;
        float ms = 0.0f;
        LBM_HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        const double mlups = static_cast<double>(b->total_cells) * 2 * pairs / (ms * 1e-3) / 1e6;
        if (mlups > best2) {
          best2 = mlups;
          b2x = tx;
          b2r = r;
          b2w = 0;
          b2nt = nt;
        }
      }
    b->tile2_x = b2x;
    b->tile2_rows = b2r;
    b->workgroups2 = b2w;
    // the stores option is shared by both kernels: the two-step one's choice when it runs
    b->store_nt = best2 > best ? b2nt : keep_nt;
    LBM_TRY(b->build_tiles2());
    if (best2 > 0.0) best = best2;
  }
  if (tile_x) *tile_x = btx;
  if (tile_rows) *tile_rows = brows;
  if (workgroups) *workgroups = bwg;
  if (best_mlups) *best_mlups = best;
  return LBM_OK;
}

}  // namespace

int lbm_batch_autotune(lbm_batch* b, int steps, int* tile_x, int* tile_rows, int* workgroups,
                       double* best_mlups) {
  // Times `steps` direct launches per candidate shape on a copy of the state, which is restored
  // afterwards whatever happens; when a trial fails, the launch configuration is restored too.
  // Every shape computes the same bits.
  if (!b || steps < 1) return fail(LBM_EINVAL, "null batch handle or steps < 1");
  const size_t bytes = b->total_elems * b->rs();
  void* backup = nullptr;
  if (hipMalloc(&backup, bytes) != hipSuccess) return fail(LBM_ENOMEM, "autotune backup allocation failed");
  const int keep_parity = b->parity;
  const int k_tx = b->tile_x, k_rows = b->tile_rows, k_wg = b->workgroups, k_order = b->order,
            k_nt = b->store_nt, k_t2x = b->tile2_x, k_t2r = b->tile2_rows, k_wg2 = b->workgroups2;
  const bool saved = hipStreamSynchronize(b->stream) == hipSuccess &&
                     hipMemcpy(backup, b->current(), bytes, hipMemcpyDeviceToDevice) == hipSuccess;
  int rc = saved ? batch_autotune_trials(b, steps, tile_x, tile_rows, workgroups, best_mlups)
                 : fail(LBM_EHIP, "autotune: saving the state failed");
  b->parity = keep_parity;
  bool restored = !saved;
  if (saved) {
    (void)hipGetLastError();  // a failed trial's launch error must not fail the restore
    restored = hipStreamSynchronize(b->stream) == hipSuccess &&
               hipMemcpy(b->current(), backup, bytes, hipMemcpyDeviceToDevice) == hipSuccess;
  }
  (void)hipFree(backup);
  if (rc != LBM_OK) {
    b->tile_x = k_tx;
    b->tile_rows = k_rows;
    b->workgroups = k_wg;
    b->order = k_order;
    b->store_nt = k_nt;
    b->tile2_x = k_t2x;
    b->tile2_rows = k_t2r;
    b->workgroups2 = k_wg2;
    // the lists of the restored shapes (a failure here is secondary to the trial's)
    if (b->build_tiles() == LBM_OK) (void)b->build_tiles2();
    return rc;
  }
  return restored ? LBM_OK : fail(LBM_EHIP, "autotune: restoring the state failed");
}

}  // extern "C"
