// Checkpointing schedules of the discrete adjoint (host only; see adjoint_schedule.hpp).
#include "adjoint_schedule.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>

namespace lbmhip {
namespace {

// ---- binomial (revolve) --------------------------------------------------------------------
// F(1, s) = 0,  F(n, 0) = n (n - 1) / 2,  F(n, s) = min_m [ m + F(n - m, s - 1) + F(m, s) ]:
// advance m steps, store R_m in a free slot, reverse the right part with one slot fewer, then
// the left part. With s >= n - 1 free slots every state is stored once.
struct Revolve {
  int n = 0, s = 0;
  std::vector<int64_t> cost;
  std::vector<int> split;
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
    if (b >= a - 1) return 1;
    return split[static_cast<size_t>(a) * (s + 1) + std::min(b, s)];
  }
};

// Emits the binomial schedule as launches. Buffers: slot k = k; two work buffers for the
// intermediate states of an advance and two "ready" buffers for states recomputed with no free
// slot (kWorkBase + 0..3).
struct BinomialEmitter {
  Revolve rv;
  AdjSchedule* out;
  int max_slot = 0;
  static constexpr int kWork0 = kWorkBase, kReady0 = kWorkBase + 2;

  void reverse(int t, int buf) {
    AdjLaunch L;
    L.rev_state = t;
    L.rev_buf = buf;
    out->launches.push_back(L);
  }
  // A forward step joins the previous launch when that launch holds only a reverse step whose
  // input it does not overwrite (the two touch disjoint buffers).
  void forward(int src, int dst, int state) {
    out->forward_steps++;
    if (!out->launches.empty()) {
      AdjLaunch& last = out->launches.back();
      if (last.rev_state >= 0 && last.nfwd == 0 && dst != last.rev_buf) {
        last.fwd[0] = {src, dst, state};
        last.nfwd = 1;
        out->fused++;
        return;
      }
    }
    AdjLaunch L;
    L.fwd[0] = {src, dst, state};
    L.nfwd = 1;
    out->launches.push_back(L);
  }
  void advance(int src, int state, int k, int dst) {
    int cur = src;
    int w = src == kWork0 ? 1 : 0;
    for (int j = 0; j < k; ++j) {
      const int o = (j == k - 1) ? dst : kWork0 + w;
      forward(cur, o, state + j);
      cur = o;
      w ^= 1;
    }
  }
  void rev(int t0, int t1, int s0, int nfree) {
    max_slot = std::max(max_slot, s0);
    const int n = t1 - t0;
    if (n == 1) return reverse(t0, s0);
    if (nfree == 0) {
      // no free slot: every R_t (t0 < t < t1) is recomputed from R_t0, alternately into the two
      // ready buffers (the recomputation of R_{t-1} never overwrites the input of reverse t)
      int b = 0;
      for (int t = t1 - 1; t > t0; --t, b ^= 1) {
        advance(s0, t0, t - t0, kReady0 + b);
        reverse(t, kReady0 + b);
      }
      return reverse(t0, s0);
    }
    const int m = rv.first_split(n, nfree);
    advance(s0, t0, m, s0 + 1);
    rev(t0 + m, t1, s0 + 1, nfree - 1);
    rev(t0, t0 + m, s0, nfree);
  }
};

// ---- pipelined multi-level -----------------------------------------------------------------
// Level-l points are the states t with t = 0 or (n - t) % K_l == 0 (segments aligned on the last
// state, so every segment but the lowest is full). The level-l segment of t is [lo, hi):
struct Levels {
  int n;
  std::vector<int> K;
  bool point(int t, int l) const { return t == 0 || (n - t) % K[l] == 0; }
  int seg_lo(int t, int l) const { return std::max(0, n - ((n - 1 - t) / K[l] + 1) * K[l]); }
  int seg_hi(int t, int l) const { return n - ((n - 1 - t) / K[l]) * K[l]; }
};

struct Chain {
  // current job: advance from state `at` (held in buffer `buf`) up to state `target`
  bool active = false;
  int at = -1, buf = -1, target = -1;
  int next_hi = -1;  // upper end of the next segment to fill (jobs run downwards)
  int work = 0;      // which of the chain's two work buffers is written next
};

bool simulate_pipelined(int n, const std::vector<int>& K, int lead, int max_slots,
                        AdjSchedule* out, std::string* why) {
  const int L = static_cast<int>(K.size());
  const int T = n - 1;
  Levels lv{n, K};
  out->launches.clear();
  out->forward_steps = 0;
  out->fused = 0;
  out->levels = L;
  out->spacing = K;

  std::vector<int> where(n, -1);  // buffer holding state t (slot), -1: not stored
  std::vector<int> free_slots;
  int nslots = 0, peak = 0, live = 0;
  auto take = [&]() {
    int s;
    if (!free_slots.empty()) {
      s = free_slots.back();
      free_slots.pop_back();
    } else {
      s = nslots++;
    }
    peak = std::max(peak, ++live);
    return s;
  };
  auto give = [&](int s) {
    free_slots.push_back(s);
    --live;
  };

  // Forward sweep: level-0 points, and for each finer level l the level-l points of the top
  // level-(l-1) segment (the first segments the reverse sweep consumes).
  auto keep = [&](int t) {
    if (lv.point(t, 0)) return true;
    for (int l = 1; l < L; ++l)
      if (lv.point(t, l) && t >= lv.seg_lo(T, l - 1)) return true;
    return false;
  };
  where[0] = take();  // slot 0: R_0
  int cur = where[0], sweep_work = 0;
  for (int t = 0; t < T; ++t) {
    AdjLaunch Lc;
    int dst;
    if (keep(t + 1)) {
      dst = take();
      where[t + 1] = dst;
    } else {
      dst = kWorkBase + sweep_work;
      sweep_work ^= 1;
    }
    Lc.fwd[0] = {cur, dst, t};
    Lc.nfwd = 1;
    out->launches.push_back(Lc);
    out->forward_steps++;
    cur = dst;
  }
  // R_T lies in the top finest segment, which the sweep keeps whole.
  if (where[T] < 0) {
    if (why) *why = "pipelined sweep did not store the last state";
    return false;
  }

  // chains l = 1 .. L-1 fill the level-l points of level-(l-1) segments, top-down, starting with
  // the segment below the top one
  std::vector<Chain> ch(L);
  for (int l = 1; l < L; ++l) ch[l].next_hi = lv.seg_lo(T, l - 1);

  int s = T;  // next state to reverse
  long long guard = 0;
  while (s >= 0) {
    if (++guard > 8LL * n + 64) {
      if (why) *why = "pipelined schedule does not terminate";
      return false;
    }
    AdjLaunch Lc;
    std::vector<std::pair<int, int>> produced;  // (state, buffer) stored after this launch
    for (int l = 1; l < L; ++l) {
      Chain& c = ch[l];
      if (!c.active) {
        // next job: the level-(l-1) segment [a, b) with b = next_hi, once the reverse sweep is
        // within `lead` + K_{l-1} states of b and R_a is stored
        while (c.next_hi > 0) {
          const int b = c.next_hi, a = lv.seg_lo(b - 1, l - 1);
          // the last level-l point below b (the finest level stores every state)
          const int target = l == L - 1 ? b - 1 : b - K[l];
          if (target <= a) {  // nothing to fill in this (short) segment
            c.next_hi = a;
            continue;
          }
          if (s < b + K[l - 1] + lead && where[a] >= 0) {
            c.active = true;
            c.at = a;
            c.buf = where[a];
            c.target = target;
            c.next_hi = a;
          }
          break;
        }
      }
      if (!c.active) continue;
      const int x = c.at;
      int dst;
      if (lv.point(x + 1, l)) {
        dst = take();
        produced.push_back({x + 1, dst});
      } else {
        dst = kWorkBase + 2 * l + c.work;
        c.work ^= 1;
      }
      Lc.fwd[Lc.nfwd++] = {c.buf, dst, x};
      out->forward_steps++;
      c.at = x + 1;
      c.buf = dst;
      if (c.at == c.target) c.active = false;
    }
    if (where[s] >= 0) {
      Lc.rev_state = s;
      Lc.rev_buf = where[s];
    }
    if (nslots > max_slots) {
      if (why) *why = "pipelined schedule exceeds the snapshot budget";
      return false;
    }
    if (Lc.rev_state < 0 && Lc.nfwd == 0) {
      if (why) *why = "pipelined schedule stalled with no runnable work";
      return false;
    }
    if (Lc.rev_state >= 0 && Lc.nfwd > 0) out->fused++;
    out->launches.push_back(Lc);
    for (auto& p : produced) where[p.first] = p.second;
    if (Lc.rev_state >= 0) {
      give(where[s]);
      where[s] = -1;
      --s;
    }
  }
  out->slots = nslots;
  (void)peak;
  out->work = 2 * L + 2;
  return true;
}

}  // namespace

bool build_binomial(int states, int slots, AdjSchedule* out, std::string* why) {
  if (states < 1 || slots < 1) {
    if (why) *why = "need states >= 1 and slots >= 1";
    return false;
  }
  slots = std::min(slots, states);
  if (slots - 1 < states - 1 &&
      std::min(slots - 1, Revolve::kMaxTableSlots) * 1.0 * states * states > 4e9) {
    if (why) *why = "schedule table too large";
    return false;
  }
  *out = AdjSchedule{};
  BinomialEmitter e;
  e.out = out;
  e.rv.build(states, slots - 1);
  e.rev(0, states, 0, slots - 1);
  out->slots = e.max_slot + 1;
  out->work = 4;
  return true;
}

namespace {
bool build_pipelined_capped(int states, const std::vector<int>& spacing, int max_slots,
                            AdjSchedule* out, std::string* why) {
  const int L = static_cast<int>(spacing.size());
  if (states < 2 || L < 2 || L > kMaxChains + 1 || spacing.back() != 1) {
    if (why) *why = "pipelined schedule needs >= 2 states and 2 .. 4 levels ending in spacing 1";
    return false;
  }
  for (int l = 0; l + 1 < L; ++l)
    if (spacing[l] <= spacing[l + 1] || spacing[l] % spacing[l + 1] != 0) {
      if (why) *why = "level spacings must decrease, each a multiple of the next";
      return false;
    }
  // the smallest lead (start of each chain's job ahead of the reverse sweep) without stalls
  // (a stall is a launch with no reverse step: the sweep then has more than 2 states launches)
  AdjSchedule best;
  bool have = false;
  for (int lead = 0; lead <= 2; ++lead) {
    AdjSchedule s;
    std::string w;
    if (!simulate_pipelined(states, spacing, lead, max_slots, &s, &w)) {
      if (why) *why = w;
      continue;
    }
    if (!have || s.launches.size() < best.launches.size() ||
        (s.launches.size() == best.launches.size() && s.slots < best.slots)) {
      best = std::move(s);
      have = true;
    }
    if (best.launches.size() <= 2 * static_cast<size_t>(states) - 1) break;  // no stall
  }
  if (!have) return false;
  *out = std::move(best);
  return true;
}
}  // namespace

bool build_pipelined(int states, const std::vector<int>& spacing, AdjSchedule* out,
                     std::string* why) {
  return build_pipelined_capped(states, spacing, INT_MAX, out, why);
}

bool build_pipelined_budget(int states, int slots, AdjSchedule* out, std::string* why) {
  if (states < 2 || slots < 2) {
    if (why) *why = "pipelined schedule needs >= 2 states and >= 2 slots";
    return false;
  }
  for (int L = 2; L <= kMaxChains + 1; ++L) {
    const double r = std::pow(static_cast<double>(states), 1.0 / L);
    const int qlo = std::max(2, static_cast<int>(r / 2)), qhi = std::max(3, static_cast<int>(2 * r) + 1);
    AdjSchedule best;
    bool have = false;
    std::vector<int> q(L - 1, qlo);
    // enumerate the multipliers q_l (K_{L-1} = 1, K_l = K_{l+1} q_l)
    for (;;) {
      std::vector<int> K(L, 1);
      for (int l = L - 2; l >= 0; --l) K[l] = K[l + 1] * q[l];
      if (K[0] < states * 2) {
        AdjSchedule s;
        if (build_pipelined_capped(states, K, slots, &s, nullptr) &&
            (!have || s.launches.size() < best.launches.size() ||
             (s.launches.size() == best.launches.size() && s.slots < best.slots))) {
          best = std::move(s);
          have = true;
        }
      }
      int d = 0;
      while (d < L - 1 && ++q[d] > qhi) q[d++] = qlo;
      if (d == L - 1) break;
    }
    if (have) {
      *out = std::move(best);
      return true;
    }
  }
  if (why) *why = "no pipelined schedule with up to 4 levels fits the snapshot budget";
  return false;
}

bool check_schedule(int states, const AdjSchedule& s, std::string* why) {
  auto bad = [&](const std::string& m) {
    if (why) *why = m;
    return false;
  };
  std::vector<std::pair<int, int>> content;  // (buffer, state)
  auto get = [&](int buf) {
    for (auto& p : content)
      if (p.first == buf) return p.second;
    return -1;
  };
  auto put = [&](int buf, int st) {
    for (auto& p : content)
      if (p.first == buf) {
        p.second = st;
        return;
      }
    content.push_back({buf, st});
  };
  put(0, 0);
  int next_rev = states - 1;
  for (size_t k = 0; k < s.launches.size(); ++k) {
    const AdjLaunch& L = s.launches[k];
    const std::string at = " (launch " + std::to_string(k) + ")";
    std::vector<int> reads, writes;
    if (L.rev_state >= 0) {
      if (L.rev_state != next_rev) return bad("reverse steps out of order" + at);
      if (get(L.rev_buf) != L.rev_state) return bad("reverse step reads the wrong state" + at);
      reads.push_back(L.rev_buf);
    }
    for (int f = 0; f < L.nfwd; ++f) {
      if (get(L.fwd[f].src) != L.fwd[f].state) return bad("forward step reads the wrong state" + at);
      if (L.fwd[f].dst < kWorkBase && L.fwd[f].dst >= s.slots) return bad("slot out of range" + at);
      reads.push_back(L.fwd[f].src);
      writes.push_back(L.fwd[f].dst);
    }
    for (int w : writes) {
      if (std::count(reads.begin(), reads.end(), w)) return bad("launch writes a buffer it reads" + at);
      if (std::count(writes.begin(), writes.end(), w) > 1) return bad("launch writes a buffer twice" + at);
    }
    for (int f = 0; f < L.nfwd; ++f) put(L.fwd[f].dst, L.fwd[f].state + 1);
    if (L.rev_state >= 0) --next_rev;
  }
  if (next_rev != -1) return bad("not every state was reversed");
  return true;
}

}  // namespace lbmhip
