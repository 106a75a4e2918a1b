// C ABI of the discrete adjoint (lbm_adjoint_*): checkpointing schedule, HIP-graph capture of
// the forward and gradient sweeps, gradient read-back. No reference counterpart.
//
// The gradient sweep is a list of launches produced by adjoint_schedule.cpp (binomial or
// pipelined multi-level checkpointing); each launch runs at most one reverse step and up to
// three forward steps (adjoint.hip). The whole sweep, including the final FP64 sums of J and
// dJ/dtau, is captured once as a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "adjoint.hpp"
#include "adjoint_schedule.hpp"
#include "d2q9.hpp"
#include "host_common.hpp"
#include "kernels.hpp"
#include "lbm_hip.h"

using namespace lbmhip;

namespace {

// Covers every cell with a nonzero link code by at most kMaxBoxes boxes: runs of consecutive
// dirty rows, each with the union of its rows' dirty column ranges; while there are too many
// runs, the two closest neighbouring runs merge (a box may cover clean cells, never the reverse).
DirtyBoxes dirty_boxes(int nx, int ny, const std::vector<uint16_t>& code) {
  struct Box { int x0, x1, y0, y1; };
  std::vector<Box> runs;
  for (int y = 0; y < ny; ++y) {
    int lo = nx, hi = -1;
    for (int x = 0; x < nx; ++x)
      if (code[static_cast<size_t>(y) * nx + x]) {
        lo = std::min(lo, x);
        hi = std::max(hi, x);
      }
    if (hi < 0) continue;
    if (!runs.empty() && runs.back().y1 == y) {
      runs.back().x0 = std::min(runs.back().x0, lo);
      runs.back().x1 = std::max(runs.back().x1, hi + 1);
      runs.back().y1 = y + 1;
    } else {
      runs.push_back({lo, hi + 1, y, y + 1});
    }
  }
  while (static_cast<int>(runs.size()) > kMaxBoxes) {
    size_t best = 0;
    int gap = ny + 1;
    for (size_t k = 0; k + 1 < runs.size(); ++k)
      if (runs[k + 1].y0 - runs[k].y1 < gap) {
        gap = runs[k + 1].y0 - runs[k].y1;
        best = k;
      }
    Box& a = runs[best];
    const Box& b = runs[best + 1];
    a = {std::min(a.x0, b.x0), std::max(a.x1, b.x1), a.y0, b.y1};
    runs.erase(runs.begin() + static_cast<long>(best) + 1);
  }
  DirtyBoxes d{};
  d.n = static_cast<int>(runs.size());
  for (int k = 0; k < d.n; ++k) {
    d.x0[k] = runs[k].x0;
    d.x1[k] = runs[k].x1;
    d.y0[k] = runs[k].y0;
    d.y1[k] = runs[k].y1;
  }
  return d;
}

}  // namespace

struct lbm_adjoint {
  int nx = 0, ny = 0, T = 0, prec = LBM_FLOAT;
  int px0 = 0, py0 = 0, pw = 1, ph = 1, avg = 1, snapshots = 1;
  int schedule_kind = LBM_SCHEDULE_AUTO;
  double tau = 0.7;
  std::vector<uint8_t> solid;
  DirtyBoxes boxes{};
  uint16_t* links = nullptr;
  void* force = nullptr;
  void* dforce = nullptr;
  // Schedule buffers: slot 0 (R_0) and work buffers 0, 1 (the forward run's) exist from creation;
  // the others are allocated with the schedule.
  std::vector<void*> slot_ptr, work_ptr;
  void* mu[2] = {nullptr, nullptr};
  double* tacc = nullptr;  // dJ/dtau accumulators, one per cell
  double* jacc = nullptr;  // J accumulators, one per probe cell
  double* sums = nullptr;  // [0] dJ/dtau, [1] J
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t fwd_exec = nullptr, grad_exec = nullptr;
  std::unique_ptr<AdjSchedule> sched;
  bool use_graph = true;

  size_t N() const { return static_cast<size_t>(nx) * ny; }
  size_t rs() const { return real_size(prec); }
  size_t state_bytes() const { return kQ * N() * rs(); }
  size_t probe_cells() const { return static_cast<size_t>(pw) * ph; }

  void drop_graphs() {
    if (fwd_exec) (void)hipGraphExecDestroy(fwd_exec);
    if (grad_exec) (void)hipGraphExecDestroy(grad_exec);
    fwd_exec = grad_exec = nullptr;
  }

  // Frees the schedule's buffers beyond the base ones (all of them with keep_base = false).
  void free_schedule_buffers(bool keep_base) {
    const size_t ks = keep_base ? 1 : 0, kw = keep_base ? 2 : 0;
    for (size_t k = ks; k < slot_ptr.size(); ++k)
      if (slot_ptr[k]) (void)hipFree(slot_ptr[k]);
    for (size_t k = kw; k < work_ptr.size(); ++k)
      if (work_ptr[k]) (void)hipFree(work_ptr[k]);
    slot_ptr.resize(std::min(slot_ptr.size(), ks));
    work_ptr.resize(std::min(work_ptr.size(), kw));
  }

  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    free_schedule_buffers(false);
    for (void* p : {static_cast<void*>(links), force, dforce, mu[0], mu[1], static_cast<void*>(tacc),
                    static_cast<void*>(jacc), static_cast<void*>(sums)})
      if (p) (void)hipFree(p);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  bool in_window(int t) const { return t > T - avg && t <= T; }

  // Schedule of the gradient sweep and its buffers (slot 0 holds R_0 = 0, the rest state).
  int ensure_schedule() {
    if (sched) return LBM_OK;
    auto s = std::make_unique<AdjSchedule>();
    std::string why;
    const int states = T + 1;
    bool ok = false;
    // auto: every state stored when the budget allows (no recomputation at all), else pipelined
    // when it fits, else binomial
    if (schedule_kind == LBM_SCHEDULE_PIPELINED ||
        (schedule_kind == LBM_SCHEDULE_AUTO && snapshots < states))
      ok = build_pipelined_budget(states, snapshots, s.get(), &why);
    if (!ok && schedule_kind == LBM_SCHEDULE_PIPELINED)
      return fail(LBM_EINVAL, "pipelined schedule: " + why);
    if (!ok && !build_binomial(states, snapshots, s.get(), &why))
      return fail(LBM_EINVAL, "binomial schedule: " + why);
    if (!check_schedule(states, *s, &why)) return fail(LBM_EHIP, "invalid schedule: " + why);
    int nwork = 2;  // the forward sweep of lbm_adjoint_forward uses work buffers 0 and 1
    for (const AdjLaunch& L : s->launches)
      for (int f = 0; f < L.nfwd; ++f)
        for (int b : {L.fwd[f].src, L.fwd[f].dst})
          if (b >= kWorkBase) nwork = std::max(nwork, b - kWorkBase + 1);
    free_schedule_buffers(true);
    while (static_cast<int>(slot_ptr.size()) < s->slots) {
      void* p = nullptr;
      if (hipMalloc(&p, state_bytes()) != hipSuccess) return fail(LBM_ENOMEM, "adjoint snapshot allocation failed");
      slot_ptr.push_back(p);
    }
    while (static_cast<int>(work_ptr.size()) < nwork) {
      void* p = nullptr;
      if (hipMalloc(&p, state_bytes()) != hipSuccess) return fail(LBM_ENOMEM, "adjoint work allocation failed");
      work_ptr.push_back(p);
    }
    sched = std::move(s);
    return LBM_OK;
  }
  void* buf(int id) const { return id >= kWorkBase ? work_ptr[id - kWorkBase] : slot_ptr[id]; }

  template <class Real>
  AdjGeom<Real> geom() const {
    AdjGeom<Real> g{};
    g.links = links;
    g.force = static_cast<const Real*>(force);
    g.N = N();
    g.nx = nx;
    g.ny = ny;
    const Real tt = static_cast<Real>(tau);
    g.omega = Real(1) / tt;
    g.fp = Real(1) - Real(0.5) * g.omega;  // as the plain step (src/v3_cuda_opt/solver_v3.cu:252-253)
    g.px0 = px0;
    g.py0 = py0;
    g.pw = pw;
    g.ph = ph;
    g.kd = 1.0 / (static_cast<double>(avg) * pw * ph);
    g.k = static_cast<Real>(g.kd);
    g.boxes = boxes;
    return g;
  }

  template <class Real>
  int launch(const AdjRev<Real>* rev, const AdjFwd<Real>* fwd, int nfwd) {
    AdjRev<Real> none{};
    const hipError_t e = launch_adjoint(geom<Real>(), rev ? *rev : none, rev != nullptr, fwd, nfwd, stream);
    if (e != hipSuccess) return fail(LBM_EHIP, std::string("adjoint launch: ") + hipGetErrorString(e));
    return LBM_OK;
  }

  template <class Real>
  int zero() {
    const hipError_t e1 = launch_fill(static_cast<Real*>(dforce), 2 * N(), Real(0), stream);
    const hipError_t e2 = launch_fill(tacc, N(), 0.0, stream);
    const hipError_t e3 = launch_fill(jacc, probe_cells(), 0.0, stream);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(LBM_EHIP, "fill launch failed");
    return LBM_OK;
  }

  // Forward run from rest: T steps through two work buffers, J accumulated inside the steps
  // whose input state lies in the window, plus a J-only pass over R_T.
  template <class Real>
  int forward_body() {
    LBM_TRY(zero<Real>());
    const void* cur = slot_ptr[0];
    for (int t = 0; t < T; ++t) {
      void* out = work_ptr[t & 1];
      const AdjFwd<Real> f{static_cast<const Real*>(cur), static_cast<Real*>(out), in_window(t) ? jacc : nullptr};
      LBM_TRY(launch<Real>(nullptr, &f, 1));
      cur = out;
    }
    if (in_window(T)) {
      const AdjFwd<Real> f{static_cast<const Real*>(cur), nullptr, jacc};
      LBM_TRY(launch<Real>(nullptr, &f, 1));
    }
    const hipError_t e = launch_sum_f64(jacc, probe_cells(), sums + 1, stream);
    if (e != hipSuccess) return fail(LBM_EHIP, "sum launch failed");
    return LBM_OK;
  }

  // Gradient: the schedule's launches; J and dJ/dtau accumulate in the reverse steps.
  template <class Real>
  int gradient_body() {
    LBM_TRY(zero<Real>());
    int cur = 0;  // mu[cur] holds mu_{t+1}
    bool have_mu = false;
    for (const AdjLaunch& L : sched->launches) {
      AdjFwd<Real> f[kAdjMaxFwd];
      for (int k = 0; k < L.nfwd; ++k)
        f[k] = {static_cast<const Real*>(buf(L.fwd[k].src)), static_cast<Real*>(buf(L.fwd[k].dst)), nullptr};
      if (L.rev_state >= 0) {
        const int t = L.rev_state;
        AdjRev<Real> r{};
        r.R = static_cast<const Real*>(buf(L.rev_buf));
        r.mu_in = have_mu ? static_cast<const Real*>(mu[cur]) : nullptr;
        r.mu_out = static_cast<Real*>(mu[cur ^ 1]);
        r.dforce = static_cast<Real*>(dforce);
        r.tacc = tacc;
        r.direct = in_window(t) ? 1 : 0;
        r.jacc = r.direct ? jacc : nullptr;
        LBM_TRY(launch<Real>(&r, f, L.nfwd));
        cur ^= 1;
        have_mu = true;
      } else {
        LBM_TRY(launch<Real>(nullptr, f, L.nfwd));
      }
    }
    const hipError_t e1 = launch_sum_f64(tacc, N(), sums, stream);
    const hipError_t e2 = launch_sum_f64(jacc, probe_cells(), sums + 1, stream);
    if (e1 != hipSuccess || e2 != hipSuccess) return fail(LBM_EHIP, "sum launch failed");
    return LBM_OK;
  }

  int body(bool gradient) {
    if (prec == LBM_DOUBLE) return gradient ? gradient_body<double>() : forward_body<double>();
    return gradient ? gradient_body<float>() : forward_body<float>();
  }

  // Captures `body` into a graph (kernels only; no memsets in the graph).
  int capture(hipGraphExec_t* exec, bool gradient) {
    hipGraph_t graph = nullptr;
    LBM_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    const int rc = body(gradient);
    const hipError_t e = hipStreamEndCapture(stream, &graph);
    if (rc != LBM_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return fail(LBM_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) return fail(LBM_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e2));
    return LBM_OK;
  }

  // Runs a sweep: replays its captured graph (built on first use), or, with
  // LBM_ADJOINT_GRAPH=0, issues the launches directly (for kernel-level profiling).
  int run(bool gradient, double* seconds) {
    if (gradient) LBM_TRY(ensure_schedule());
    hipGraphExec_t* exec = gradient ? &grad_exec : &fwd_exec;
    if (use_graph && !*exec) LBM_TRY(capture(exec, gradient));
    LBM_HIP_TRY(hipEventRecord(ev0, stream));
    if (use_graph)
      LBM_HIP_TRY(hipGraphLaunch(*exec, stream));
    else
      LBM_TRY(body(gradient));
    LBM_HIP_TRY(hipEventRecord(ev1, stream));
    LBM_HIP_TRY(hipEventSynchronize(ev1));
    float ms = 0.0f;
    LBM_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    if (seconds) *seconds = ms * 1e-3;
    return LBM_OK;
  }
};

extern "C" {

int lbm_adjoint_create(int nx, int ny, double tau, const uint8_t* solid, int steps, int probe_x0,
                       int probe_y0, int probe_w, int probe_h, int avg_steps, int snapshots,
                       int precision, lbm_adjoint** out) {
  if (!out) return fail(LBM_EINVAL, "out handle pointer is null");
  *out = nullptr;
  if (nx < 1 || ny < 1 || !solid || steps < 1) return fail(LBM_EINVAL, "invalid grid, mask or steps");
  if (precision != LBM_FLOAT && precision != LBM_DOUBLE)
    return fail(LBM_EINVAL, "the adjoint runs in LBM_FLOAT or LBM_DOUBLE");
  if (static_cast<uint64_t>(kQ) * nx * ny * real_size(precision) >= (uint64_t{1} << 32))
    return fail(LBM_EINVAL, "the adjoint's states must stay below 4 GiB");
  if (probe_w < 1 || probe_h < 1 || probe_x0 < 0 || probe_y0 < 0 || probe_x0 + probe_w > nx ||
      probe_y0 + probe_h > ny || probe_w * probe_h > (1 << 20))
    return fail(LBM_EINVAL, "probe window outside the grid");
  if (avg_steps < 1 || avg_steps > steps + 1) return fail(LBM_EINVAL, "avg_steps must be in 1..steps+1");
  if (snapshots < 1) return fail(LBM_EINVAL, "need at least one snapshot slot");
  const bool ok = precision == LBM_DOUBLE ? tau > 0.5 : static_cast<float>(tau) > 0.5f;
  if (!ok) return fail(LBM_EINVAL, "tau must be > 0.5");
  auto* a = new lbm_adjoint();
  auto bail = [&](int rc) {
    a->release();
    delete a;
    return rc;
  };
  a->nx = nx;
  a->ny = ny;
  a->T = steps;
  a->prec = precision;
  a->tau = tau;
  a->px0 = probe_x0;
  a->py0 = probe_y0;
  a->pw = probe_w;
  a->ph = probe_h;
  a->avg = avg_steps;
  a->snapshots = std::min(snapshots, steps + 1);
  const char* genv = std::getenv("LBM_ADJOINT_GRAPH");
  a->use_graph = !(genv && std::string(genv) == "0");
  a->solid.assign(solid, solid + a->N());
  const Links L = build_links(nx, ny, a->solid.data(), false, true);
  a->boxes = dirty_boxes(nx, ny, L.code);
  if (hipStreamCreate(&a->stream) != hipSuccess || hipEventCreate(&a->ev0) != hipSuccess ||
      hipEventCreate(&a->ev1) != hipSuccess)
    return bail(fail(LBM_EHIP, "stream / event creation failed"));
  const size_t sb = a->state_bytes();
  if (hipMalloc(&a->links, a->N() * sizeof(uint16_t)) != hipSuccess ||
      hipMalloc(&a->force, 2 * a->N() * a->rs()) != hipSuccess ||
      hipMalloc(&a->dforce, 2 * a->N() * a->rs()) != hipSuccess ||
      hipMalloc(&a->mu[0], sb) != hipSuccess || hipMalloc(&a->mu[1], sb) != hipSuccess ||
      hipMalloc(&a->tacc, a->N() * sizeof(double)) != hipSuccess ||
      hipMalloc(&a->jacc, a->probe_cells() * sizeof(double)) != hipSuccess ||
      hipMalloc(&a->sums, 2 * sizeof(double)) != hipSuccess)
    return bail(fail(LBM_ENOMEM, "device allocation of the adjoint failed"));
  for (int k = 0; k < 3; ++k) {
    void* p = nullptr;
    if (hipMalloc(&p, sb) != hipSuccess) return bail(fail(LBM_ENOMEM, "device allocation of the adjoint failed"));
    (k == 0 ? a->slot_ptr : a->work_ptr).push_back(p);
  }
  if (hipMemcpy(a->links, L.code.data(), L.code.size() * sizeof(uint16_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(a->force, 0, 2 * a->N() * a->rs()) != hipSuccess ||
      hipMemset(a->slot_ptr[0], 0, sb) != hipSuccess)  // rest initial state: g = f - w = 0
    return bail(fail(LBM_EHIP, "adjoint upload failed"));
  *out = a;
  return LBM_OK;
}

int lbm_adjoint_set_schedule(lbm_adjoint* a, int kind) {
  if (!a) return fail(LBM_EINVAL, "null adjoint handle");
  if (kind != LBM_SCHEDULE_AUTO && kind != LBM_SCHEDULE_BINOMIAL && kind != LBM_SCHEDULE_PIPELINED)
    return fail(LBM_EINVAL, "unknown schedule kind");
  LBM_HIP_TRY(hipStreamSynchronize(a->stream));
  a->schedule_kind = kind;
  a->sched.reset();
  a->drop_graphs();  // buffer addresses are captured in the graphs
  return a->ensure_schedule();
}

int lbm_adjoint_schedule_info(lbm_adjoint* a, int* kind, int* slots, int* levels,
                              long long* forward_steps, long long* launches) {
  if (!a) return fail(LBM_EINVAL, "null adjoint handle");
  LBM_TRY(a->ensure_schedule());
  if (kind) *kind = a->sched->levels ? LBM_SCHEDULE_PIPELINED : LBM_SCHEDULE_BINOMIAL;
  if (slots) *slots = a->sched->slots;
  if (levels) *levels = a->sched->levels;
  if (forward_steps) *forward_steps = a->sched->forward_steps;
  if (launches) *launches = static_cast<long long>(a->sched->launches.size());
  return LBM_OK;
}

int lbm_checkpoint_schedule(int states, int slots, int kind, int* slots_used, int* levels,
                            long long* forward_steps, long long* launches) {
  if (states < 1 || slots < 1) return fail(LBM_EINVAL, "need states >= 1 and slots >= 1");
  AdjSchedule s;
  std::string why;
  bool ok = false;
  if (kind == LBM_SCHEDULE_PIPELINED || (kind == LBM_SCHEDULE_AUTO && slots < states))
    ok = states >= 2 && build_pipelined_budget(states, slots, &s, &why);
  if (!ok && kind == LBM_SCHEDULE_PIPELINED) return fail(LBM_EINVAL, "pipelined schedule: " + why);
  if (!ok && kind != LBM_SCHEDULE_AUTO && kind != LBM_SCHEDULE_BINOMIAL)
    return fail(LBM_EINVAL, "unknown schedule kind");
  if (!ok && !build_binomial(states, slots, &s, &why)) return fail(LBM_EINVAL, "binomial schedule: " + why);
  if (!check_schedule(states, s, &why)) return fail(LBM_EINVAL, "invalid schedule: " + why);
  if (slots_used) *slots_used = s.slots;
  if (levels) *levels = s.levels;
  if (forward_steps) *forward_steps = s.forward_steps;
  if (launches) *launches = static_cast<long long>(s.launches.size());
  return LBM_OK;
}

int lbm_checkpoint_forward_steps(int states, int slots, long long* steps) {
  if (!steps) return fail(LBM_EINVAL, "null output");
  long long f = 0;
  LBM_TRY(lbm_checkpoint_schedule(states, slots, LBM_SCHEDULE_BINOMIAL, nullptr, nullptr, &f, nullptr));
  *steps = f;
  return LBM_OK;
}

int lbm_adjoint_destroy(lbm_adjoint* a) {
  if (!a) return LBM_OK;
  a->release();
  delete a;
  return LBM_OK;
}

int lbm_adjoint_set_force_field(lbm_adjoint* a, const void* fxy) {
  if (!a || !fxy) return fail(LBM_EINVAL, "null adjoint handle or field");
  LBM_HIP_TRY(hipStreamSynchronize(a->stream));
  LBM_HIP_TRY(hipMemcpy(a->force, fxy, 2 * a->N() * a->rs(), hipMemcpyHostToDevice));
  return LBM_OK;
}

int lbm_adjoint_set_tau(lbm_adjoint* a, double tau) {
  if (!a) return fail(LBM_EINVAL, "null adjoint handle");
  const bool ok = a->prec == LBM_DOUBLE ? tau > 0.5 : static_cast<float>(tau) > 0.5f;
  if (!ok) return fail(LBM_EINVAL, "tau must be > 0.5");
  LBM_HIP_TRY(hipStreamSynchronize(a->stream));
  a->tau = tau;
  a->drop_graphs();  // omega is a kernel argument captured in the graphs
  return LBM_OK;
}

int lbm_adjoint_forward(lbm_adjoint* a, double* J, double* seconds) {
  if (!a) return fail(LBM_EINVAL, "null adjoint handle");
  LBM_TRY(a->run(false, seconds));
  if (J) LBM_HIP_TRY(hipMemcpy(J, a->sums + 1, sizeof(double), hipMemcpyDeviceToHost));
  return LBM_OK;
}

int lbm_adjoint_gradient(lbm_adjoint* a, double* J, void* dJ_dforce, double* dJ_dtau,
                         double* seconds, long long* forward_steps) {
  if (!a) return fail(LBM_EINVAL, "null adjoint handle");
  LBM_TRY(a->run(true, seconds));
  if (dJ_dforce)
    LBM_HIP_TRY(hipMemcpy(dJ_dforce, a->dforce, 2 * a->N() * a->rs(), hipMemcpyDeviceToHost));
  double s[2] = {0.0, 0.0};
  LBM_HIP_TRY(hipMemcpy(s, a->sums, sizeof(s), hipMemcpyDeviceToHost));
  if (dJ_dtau) *dJ_dtau = s[0];
  if (J) *J = s[1];
  if (forward_steps) *forward_steps = a->sched->forward_steps;
  return LBM_OK;
}

}  // extern "C"
