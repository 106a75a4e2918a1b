// C ABI of the discrete adjoint (lbm_adjoint_*): binomial checkpointing schedule, HIP-graph
// capture of the forward and reverse sweeps, gradient read-back. No reference counterpart.
//
// Checkpointing: the reverse sweep needs R_t for t = T, T-1, ..., 0. With s snapshot slots
// (slot 0 holds R_0), the schedule minimises the forward steps recomputed, by the binomial
// (revolve) recursion
//   F(1, s) = 0,  F(n, 0) = n (n - 1) / 2,
//   F(n, s) = min_m [ m + F(n - m, s - 1) + F(m, s) ]
// (advance m steps, store R_m in a free slot, reverse the right part with one slot fewer, then
// the left part). With s >= T + 1 slots every state is stored once and nothing is recomputed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstdint>
#include <string>
#include <vector>

#include "adjoint.hpp"
#include "d2q9.hpp"
#include "host_common.hpp"
#include "kernels.hpp"
#include "lbm_hip.h"

using namespace lbmhip;

namespace {

// Optimal first split of n states with s free slots (binomial checkpointing).
struct Revolve {
  int n = 0, s = 0;
  std::vector<int64_t> cost;  // (n + 1) x (s + 1)
  std::vector<int> split;

  // With s_ >= n_ - 1 free slots every state is stored (no table needed). Otherwise the table
  // is O(n^2 s); it is built for at most kMaxTableSlots slots (more slots than that and fewer
  // than n - 1 use the first kMaxTableSlots only: still a valid, slightly costlier schedule).
  static constexpr int kMaxTableSlots = 128;
  void build(int n_, int s_) {
    n = n_;
    if (s_ >= n_ - 1) {
      s = s_;
      return;
    }
    s = std::min(s_, kMaxTableSlots);
    cost.assign(static_cast<size_t>(n + 1) * (s + 1), 0);
    split.assign(static_cast<size_t>(n + 1) * (s + 1), 1);
    auto C = [&](int a, int b) -> int64_t& { return cost[static_cast<size_t>(a) * (s + 1) + b]; };
    auto M = [&](int a, int b) -> int& { return split[static_cast<size_t>(a) * (s + 1) + b]; };
    for (int a = 1; a <= n; ++a) C(a, 0) = static_cast<int64_t>(a) * (a - 1) / 2;
    for (int b = 1; b <= s; ++b)
      for (int a = 2; a <= n; ++a) {
        int64_t best = INT64_MAX;
        int bm = 1;
        for (int m = 1; m < a; ++m) {
          const int64_t c = m + C(a - m, b - 1) + C(m, b);
          if (c < best) {
            best = c;
            bm = m;
          }
        }
        C(a, b) = best;
        M(a, b) = bm;
      }
  }
  int first_split(int a, int b) const {
    if (b >= a - 1) return 1;  // enough slots: store every state
    return split[static_cast<size_t>(a) * (s + 1) + std::min(b, s)];
  }
  int64_t forward_steps(int a, int b) const {
    if (b >= a - 1) return a - 1;
    return cost[static_cast<size_t>(a) * (s + 1) + std::min(b, s)];
  }
};

}  // namespace

struct lbm_adjoint {
  int nx = 0, ny = 0, T = 0, prec = LBM_FLOAT;
  int px0 = 0, py0 = 0, pw = 1, ph = 1, avg = 1, snapshots = 1;
  double tau = 0.7;
  std::vector<uint8_t> solid;
  uint16_t* links = nullptr;
  uint8_t* segflags = nullptr;
  int nseg = 0;
  void* force = nullptr;
  void* dforce = nullptr;
  void* slots = nullptr;
  void* work[2] = {nullptr, nullptr};   // intermediate states of a recomputation
  void* ready[2] = {nullptr, nullptr};  // recomputed R_t waiting for its reverse step
  void* mu[2] = {nullptr, nullptr};
  double* partials = nullptr;  // per reverse-kernel block: dJ/dtau, J
  double* jacc = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t fwd_exec = nullptr, grad_exec = nullptr;
  Revolve rv;
  long long forward_steps_scheduled = 0;
  int mu_cur = 0;  // index of the mu buffer holding mu_{t+1}
  bool fuse = true;       // reverse steps share launches with forward steps (see forward_step)
  bool use_graph = true;  // sweeps replayed as captured HIP graphs
  bool have_mu = false;

  size_t N() const { return static_cast<size_t>(nx) * ny; }
  size_t rs() const { return real_size(prec); }
  size_t state_bytes() const { return kQ * N() * rs(); }
  void* slot(int k) const { return static_cast<char*>(slots) + static_cast<size_t>(k) * state_bytes(); }

  void drop_graphs() {
    if (fwd_exec) (void)hipGraphExecDestroy(fwd_exec);
    if (grad_exec) (void)hipGraphExecDestroy(grad_exec);
    fwd_exec = grad_exec = nullptr;
  }

  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    for (void* p : {static_cast<void*>(links), static_cast<void*>(segflags), force, dforce, slots,
                    work[0], work[1], ready[0], ready[1], mu[0], mu[1], static_cast<void*>(partials),
                    static_cast<void*>(jacc)})
      if (p) (void)hipFree(p);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  template <class T>
  StepParams<T> step_params(const void* src, void* dst, int block_x, int rows) const {
    StepParams<T> p{};
    p.g = static_cast<const T*>(src);
    p.gn = static_cast<T*>(dst);
    p.links = links;
    p.segflags = segflags;
    p.force = static_cast<const T*>(force);
    p.N = N();
    p.nx = nx;
    p.ny = ny;
    p.row0 = 0;
    p.nrows = ny;
    p.rows_per_tile = rows;
    p.tiles_x = (nx + block_x - 1) / block_x;
    p.nseg = nseg;
    p.wrap_y = 1;
    p.one_rsrc = 1;
    p.xcd_bands = 1;  // measured faster than round-robin on this small, latency-bound grid
    const T t = static_cast<T>(tau);
    p.omega = T(1) / t;
    p.fp = T(1) - T(0.5) * p.omega;
    return p;
  }

  template <class T>
  int step(const void* src, void* dst) {
    const hipError_t e = launch_step(step_params<T>(src, dst, 64, 1), false, 64, stream);
    if (e != hipSuccess) return fail(LBM_EHIP, std::string("adjoint forward: ") + hipGetErrorString(e));
    return LBM_OK;
  }
  int step_any(const void* src, void* dst) {
    return prec == LBM_DOUBLE ? step<double>(src, dst) : step<float>(src, dst);
  }

  bool in_window(int t) const { return t > T - avg && t <= T; }

  template <class T>
  int probe(const void* R) {
    GatherParams<T> p{};
    p.g = static_cast<const T*>(R);
    p.links = links;
    p.force = static_cast<const T*>(force);
    p.N = N();
    p.nx = nx;
    p.ny = ny;
    const double k = 1.0 / (static_cast<double>(avg) * pw * ph);
    const hipError_t e = launch_probe(p, px0, py0, pw, ph, k, jacc, stream);
    if (e != hipSuccess) return fail(LBM_EHIP, std::string("probe: ") + hipGetErrorString(e));
    return LBM_OK;
  }
  int probe_any(const void* R) { return prec == LBM_DOUBLE ? probe<double>(R) : probe<float>(R); }

  template <class T>
  AdjointParams<T> reverse_params(int t, const void* R) const {
    AdjointParams<T> p{};
    p.R = static_cast<const T*>(R);
    p.mu_in = have_mu ? static_cast<const T*>(mu[mu_cur]) : nullptr;
    p.mu_out = static_cast<T*>(mu[mu_cur ^ 1]);
    p.links = links;
    p.force = static_cast<const T*>(force);
    p.dforce = static_cast<T*>(dforce);
    p.partials = partials;
    p.N = N();
    p.nx = nx;
    p.ny = ny;
    const T tt = static_cast<T>(tau);
    p.omega = T(1) / tt;
    p.fp = T(1) - T(0.5) * p.omega;
    p.direct = in_window(t) ? 1 : 0;
    p.px0 = px0;
    p.py0 = py0;
    p.pw = pw;
    p.ph = ph;
    p.kd = 1.0 / (static_cast<double>(avg) * pw * ph);
    p.k = static_cast<T>(p.kd);
    return p;
  }

  template <class T>
  int reverse(int t, const void* R) {
    const hipError_t e = launch_adjoint_step(reverse_params<T>(t, R), stream);
    if (e != hipSuccess) return fail(LBM_EHIP, std::string("adjoint step: ") + hipGetErrorString(e));
    mu_cur ^= 1;
    have_mu = true;
    return LBM_OK;
  }
  int reverse_any(int t, const void* R) {
    return prec == LBM_DOUBLE ? reverse<double>(t, R) : reverse<float>(t, R);
  }

  // Reverse step t on the state held in R and, in the same launch, one forward step src -> dst.
  template <class T>
  int fused(int t, const void* R, const void* src, void* dst) {
    const hipError_t e = launch_adjoint_fused(reverse_params<T>(t, R),
                                              step_params<T>(src, dst, kAdjointThreads, 1), stream);
    if (e != hipSuccess) return fail(LBM_EHIP, std::string("fused adjoint step: ") + hipGetErrorString(e));
    mu_cur ^= 1;
    have_mu = true;
    return LBM_OK;
  }
  int fused_any(int t, const void* R, const void* src, void* dst) {
    return prec == LBM_DOUBLE ? fused<double>(t, R, src, dst) : fused<float>(t, R, src, dst);
  }

  // A reverse step is issued lazily: it waits for the next operation and, when that is a
  // forward step that does not overwrite the reverse step's input state, both share one launch
  // (they touch disjoint buffers: the reverse step reads R_t and mu_{t+1} and writes mu_t and
  // the gradient accumulators, which no forward step reads). Reverse steps keep their order.
  struct Pending {
    bool on = false;
    int t = 0;
    const void* R = nullptr;
  } pending;
  long long fused_launches = 0;

  int flush() {
    if (!pending.on) return LBM_OK;
    pending.on = false;
    return reverse_any(pending.t, pending.R);
  }
  int queue_reverse(int t, const void* R) {
    LBM_TRY(flush());
    pending = {true, t, R};
    return LBM_OK;
  }
  int forward_step(const void* src, void* dst) {
    forward_steps_scheduled++;
    if (pending.on && fuse && dst != pending.R) {
      pending.on = false;
      fused_launches++;
      return fused_any(pending.t, pending.R, src, dst);
    }
    LBM_TRY(flush());
    return step_any(src, dst);
  }

  // Advances from the state held in `src` by k >= 1 steps; the last step writes into `dst`,
  // the others into the work buffers (never the one `src` may be).
  int advance(const void* src, int k, void* dst) {
    const void* cur = src;
    int w = src == work[0] ? 1 : 0;
    for (int j = 0; j < k; ++j) {
      void* out = (j == k - 1) ? dst : work[w];
      LBM_TRY(forward_step(cur, out));
      cur = out;
      w ^= 1;
    }
    return LBM_OK;
  }

  // Reverse ops for the states t1 - 1, ..., t0; R_{t0} is held in slot s0; slots s0 + 1 ...
  // s0 + nfree are free.
  int rev(int t0, int t1, int s0, int nfree) {
    const int n = t1 - t0;
    if (n == 1) return queue_reverse(t0, slot(s0));
    if (nfree == 0) {
      // No free slot: every R_t (t0 < t < t1) is recomputed from R_{t0}, alternately into the
      // two ready buffers, so the recomputation of R_{t-1} never overwrites the input of the
      // still-pending reverse step t.
      int b = 0;
      for (int t = t1 - 1; t > t0; --t, b ^= 1) {
        LBM_TRY(advance(slot(s0), t - t0, ready[b]));
        LBM_TRY(queue_reverse(t, ready[b]));
      }
      return queue_reverse(t0, slot(s0));
    }
    const int m = rv.first_split(n, nfree);
    LBM_TRY(advance(slot(s0), m, slot(s0 + 1)));
    LBM_TRY(rev(t0 + m, t1, s0 + 1, nfree - 1));
    return rev(t0, t0 + m, s0, nfree);
  }

  template <class T>
  int fills() {
    const hipError_t e1 = launch_fill(static_cast<T*>(dforce), 2 * N(), T(0), stream);
    const hipError_t e2 = launch_fill(partials, 2 * static_cast<size_t>(adjoint_blocks(N())), 0.0, stream);
    const hipError_t e3 = launch_fill(jacc, 1, 0.0, stream);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(LBM_EHIP, "fill launch failed");
    return LBM_OK;
  }

  // Captures `body` into a graph (kernels only; no memsets in the graph).
  template <class F>
  int capture(hipGraphExec_t* exec, F&& body) {
    hipGraph_t graph = nullptr;
    LBM_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    const int rc = body();
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

  int forward_body() {
    LBM_TRY(prec == LBM_DOUBLE ? fills<double>() : fills<float>());
    const void* cur = slot(0);
    if (in_window(0)) LBM_TRY(probe_any(cur));
    for (int t = 0; t < T; ++t) {
      void* out = work[t & 1];
      LBM_TRY(step_any(cur, out));
      cur = out;
      if (in_window(t + 1)) LBM_TRY(probe_any(cur));
    }
    return LBM_OK;
  }

  int gradient_body() {
    forward_steps_scheduled = 0;
    LBM_TRY(prec == LBM_DOUBLE ? fills<double>() : fills<float>());
    mu_cur = 0;
    have_mu = false;
    pending.on = false;
    fused_launches = 0;
    LBM_TRY(rev(0, T + 1, 0, snapshots - 1));
    return flush();
  }

  // Runs a sweep: replays its captured graph (built on first use), or, with
  // LBM_ADJOINT_GRAPH=0, issues the launches directly (for kernel-level profiling).
  int run(bool gradient, double* seconds) {
    hipGraphExec_t* exec = gradient ? &grad_exec : &fwd_exec;
    if (use_graph && !*exec)
      LBM_TRY(capture(exec, [&]() -> int { return gradient ? gradient_body() : forward_body(); }));
    LBM_HIP_TRY(hipEventRecord(ev0, stream));
    if (use_graph)
      LBM_HIP_TRY(hipGraphLaunch(*exec, stream));
    else
      LBM_TRY(gradient ? gradient_body() : forward_body());
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
  // The fused launch addresses the nine planes through one buffer resource (< 4 GiB);
  // LBM_ADJOINT_FUSE=0 turns it off (separate launches, same bits).
  const char* fenv = std::getenv("LBM_ADJOINT_FUSE");
  a->fuse = kQ * a->N() * real_size(precision) < (size_t(1) << 32) &&
            !(fenv && std::string(fenv) == "0");
  const char* genv = std::getenv("LBM_ADJOINT_GRAPH");
  a->use_graph = !(genv && std::string(genv) == "0");
  a->solid.assign(solid, solid + a->N());
  a->nseg = (nx + 63) / 64;
  if (hipStreamCreate(&a->stream) != hipSuccess || hipEventCreate(&a->ev0) != hipSuccess ||
      hipEventCreate(&a->ev1) != hipSuccess)
    return bail(fail(LBM_EHIP, "stream / event creation failed"));
  const size_t sb = a->state_bytes();
  if (hipMalloc(&a->links, a->N() * sizeof(uint16_t)) != hipSuccess ||
      hipMalloc(&a->segflags, static_cast<size_t>(ny) * a->nseg) != hipSuccess ||
      hipMalloc(&a->force, 2 * a->N() * a->rs()) != hipSuccess ||
      hipMalloc(&a->dforce, 2 * a->N() * a->rs()) != hipSuccess ||
      hipMalloc(&a->slots, sb * a->snapshots) != hipSuccess || hipMalloc(&a->work[0], sb) != hipSuccess ||
      hipMalloc(&a->work[1], sb) != hipSuccess || hipMalloc(&a->ready[0], sb) != hipSuccess ||
      hipMalloc(&a->ready[1], sb) != hipSuccess || hipMalloc(&a->mu[0], sb) != hipSuccess ||
      hipMalloc(&a->mu[1], sb) != hipSuccess ||
      hipMalloc(&a->partials, 2 * sizeof(double) * adjoint_blocks(a->N())) != hipSuccess ||
      hipMalloc(&a->jacc, sizeof(double)) != hipSuccess)
    return bail(fail(LBM_ENOMEM, "device allocation of the adjoint failed"));
  const Links L = build_links(nx, ny, a->solid.data(), false, true);
  if (hipMemcpy(a->links, L.code.data(), L.code.size() * sizeof(uint16_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(a->segflags, L.seg.data(), L.seg.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(a->force, 0, 2 * a->N() * a->rs()) != hipSuccess ||
      hipMemset(a->slot(0), 0, sb) != hipSuccess)  // rest initial state: g = f - w = 0
    return bail(fail(LBM_EHIP, "adjoint upload failed"));
  a->rv.build(a->T + 1, a->snapshots - 1);
  *out = a;
  return LBM_OK;
}

int lbm_checkpoint_forward_steps(int states, int slots, long long* steps) {
  if (states < 1 || slots < 1 || !steps) return fail(LBM_EINVAL, "need states >= 1, slots >= 1");
  if (slots - 1 < states - 1 && std::min(slots - 1, Revolve::kMaxTableSlots) * 1.0 * states * states > 4e9)
    return fail(LBM_EINVAL, "schedule table too large");
  Revolve r;
  r.build(states, slots - 1);
  *steps = r.forward_steps(states, slots - 1);
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
  if (J) LBM_HIP_TRY(hipMemcpy(J, a->jacc, sizeof(double), hipMemcpyDeviceToHost));
  return LBM_OK;
}

int lbm_adjoint_gradient(lbm_adjoint* a, double* J, void* dJ_dforce, double* dJ_dtau,
                         double* seconds, long long* forward_steps) {
  if (!a) return fail(LBM_EINVAL, "null adjoint handle");
  LBM_TRY(a->run(true, seconds));
  if (dJ_dforce)
    LBM_HIP_TRY(hipMemcpy(dJ_dforce, a->dforce, 2 * a->N() * a->rs(), hipMemcpyDeviceToHost));
  std::vector<double> part(2 * static_cast<size_t>(adjoint_blocks(a->N())));
  LBM_HIP_TRY(hipMemcpy(part.data(), a->partials, part.size() * sizeof(double), hipMemcpyDeviceToHost));
  double st = 0.0, sj = 0.0;  // fixed order: deterministic
  for (size_t b = 0; b < part.size() / 2; ++b) {
    st += part[2 * b];
    sj += part[2 * b + 1];
  }
  if (dJ_dtau) *dJ_dtau = st;
  if (J) *J = sj;
  if (forward_steps) *forward_steps = a->forward_steps_scheduled;
  return LBM_OK;
}

}  // extern "C"
