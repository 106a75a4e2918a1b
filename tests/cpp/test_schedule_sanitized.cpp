// The adjoint's checkpointing schedules (host code of the library, csrc/adjoint_schedule.cpp)
// built with AddressSanitizer and UndefinedBehaviorSanitizer: every builder over a sweep of
// (states, slots) pairs and level spacings, each schedule replayed on buffer contents by
// check_schedule(), and the invariants the GPU sweeps rely on checked here (reverse launches =
// states, forward steps of the first sweep included, the stored-state peak within the budget,
// chains within kMaxChains). Any heap, stack or integer fault aborts the program (the sanitizers
// run with halt_on_error); the checks themselves print the first failing case and return 1.
#include <cstdio>
#include <string>
#include <vector>

#include "adjoint_schedule.hpp"

using lbmhip::AdjSchedule;

namespace {

int failures = 0;

void fail(const char* what, int states, int slots, const std::string& why) {
  if (failures++ < 10) std::printf("FAIL %s states=%d slots=%d: %s\n", what, states, slots, why.c_str());
}

// Properties every schedule must have beyond check_schedule's replay.
void check_common(const char* what, int states, int slots, const AdjSchedule& s) {
  std::string why;
  if (!lbmhip::check_schedule(states, s, &why)) return fail(what, states, slots, why);
  long long rev = 0, fwd = 0;
  for (const auto& L : s.launches) {
    if (L.rev_state >= 0) ++rev;
    if (L.nfwd < 0 || L.nfwd > lbmhip::kMaxChains) return fail(what, states, slots, "chain count");
    fwd += L.nfwd;
    if (L.rev_state < 0 && L.nfwd == 0) return fail(what, states, slots, "empty launch");
  }
  if (rev != states) return fail(what, states, slots, "reverse launches != states");
  if (fwd != s.forward_steps) return fail(what, states, slots, "forward step count");
  if (fwd < states - 1) return fail(what, states, slots, "fewer forward steps than one sweep");
}

}  // namespace

int main() {
  long long built = 0;
  for (int states = 1; states <= 260; states += (states < 40 ? 1 : 7)) {
    for (int slots = 1; slots <= 48; slots += (slots < 12 ? 1 : 5)) {
      AdjSchedule b;
      std::string why;
      if (!lbmhip::build_binomial(states, slots, &b, &why)) {
        fail("binomial", states, slots, why);
        continue;
      }
      ++built;
      check_common("binomial", states, slots, b);
      if (b.slots > slots) fail("binomial", states, slots, "more stored states than the budget");
      if (states >= 2 && slots >= 2) {
        AdjSchedule p;
        if (lbmhip::build_pipelined_budget(states, slots, &p, &why)) {
          ++built;
          check_common("pipelined budget", states, slots, p);
          if (p.slots > slots) fail("pipelined budget", states, slots, "over budget");
          if (p.forward_steps < b.forward_steps)
            fail("pipelined budget", states, slots, "fewer forward steps than the binomial optimum");
        }
      }
    }
  }
  // explicit spacings, two to four levels, states not multiples of the spacings included
  const std::vector<std::vector<int>> spacings = {{2, 1}, {5, 1}, {12, 3, 1}, {30, 6, 1},
                                                  {64, 16, 4, 1}, {143, 11, 1}};
  for (int states = 2; states <= 2001; states += (states < 100 ? 3 : 211)) {
    for (const auto& sp : spacings) {
      AdjSchedule p;
      std::string why;
      if (!lbmhip::build_pipelined(states, sp, &p, &why)) continue;  // e.g. spacing beyond states
      ++built;
      check_common("pipelined", states, sp[0], p);
    }
  }
  // the differentiation instance: T = 2000, 45 snapshots, both families
  {
    AdjSchedule b, p;
    std::string why;
    if (!lbmhip::build_binomial(2001, 45, &b, &why)) fail("binomial", 2001, 45, why);
    else check_common("binomial", 2001, 45, b);
    if (!lbmhip::build_pipelined_budget(2001, 45, &p, &why)) fail("pipelined budget", 2001, 45, why);
    else check_common("pipelined budget", 2001, 45, p);
    built += 2;
  }
  // negative controls: the replay must reject a schedule with a planted mistake
  {
    AdjSchedule b;
    std::string why;
    if (lbmhip::build_binomial(40, 6, &b, &why)) {
      AdjSchedule bad = b;
      for (auto& L : bad.launches)
        if (L.rev_state == 20) {
          L.rev_state = 19;  // reverses a state out of order
          break;
        }
      if (lbmhip::check_schedule(40, bad, &why)) fail("planted reverse-order mistake accepted", 40, 6, "");
      bad = b;
      for (auto& L : bad.launches)
        if (L.nfwd > 0 && L.rev_state < 0 && L.fwd[0].state > 5) {
          L.fwd[0].state -= 1;  // a forward step that reads a state its source buffer does not hold
          break;
        }
      if (lbmhip::check_schedule(40, bad, &why)) fail("planted forward-step mistake accepted", 40, 6, "");
    } else {
      fail("binomial", 40, 6, why);
    }
  }
  // invalid arguments are refused, not faulted on
  {
    AdjSchedule s;
    std::string why;
    if (lbmhip::build_binomial(0, 4, &s, &why) || lbmhip::build_binomial(5, 0, &s, &why) ||
        lbmhip::build_pipelined(5, {3, 2}, &s, &why) || lbmhip::build_pipelined(5, {1}, &s, &why) ||
        lbmhip::build_pipelined_budget(1, 4, &s, &why))
      fail("invalid arguments accepted", 0, 0, "");
  }
  std::printf("%lld schedules built and replayed, %d failures\n", built, failures);
  return failures ? 1 : 0;
}
