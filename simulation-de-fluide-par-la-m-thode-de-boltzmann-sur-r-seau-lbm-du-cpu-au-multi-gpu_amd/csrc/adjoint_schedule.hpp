// Checkpointing schedules of the discrete adjoint, as a list of launches (host only).
//
// The reverse sweep needs the forward states R_t for t = T, T-1, ..., 0 (one reverse step per
// state, in that order). A schedule says, launch by launch, which state is reversed and which
// forward steps (recomputations) run in the same launch, and in which buffers every state
// lives. Two families:
//
//   binomial   (revolve) the fewest recomputed forward steps for a number of stored states
//              (Griewank's binomial schedule, by exact dynamic programming); forward steps run
//              one after another, a reverse step shares a launch with a following forward step
//              only where the schedule happens to allow it.
//   pipelined  multi-level checkpoints (level-0 points every K0 states stored by the forward
//              sweep, level-1 points every K1 inside one level-0 segment, ..., every state inside
//              one finest segment); each level has one forward chain that recomputes the points
//              of the next segment to be reversed while the current one is reversed, so every
//              launch holds one reverse step plus one forward step per chain. The reverse
//              sweep then costs about T launches of (reverse + chains) instead of T reverse
//              launches plus about (r - 1) T recomputation launches.
//
// Buffers: ids [0, slots) are snapshot slots (slot 0 holds R_0), ids kWorkBase + k are work
// buffers (intermediate states of a forward chain, never kept).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace lbmhip {

constexpr int kWorkBase = 1 << 20;
constexpr int kMaxChains = 3;

struct AdjFwdOp {
  int src = -1, dst = -1;  // buffer ids
  int state = -1;          // the step computes R_{state + 1} from R_state
};

struct AdjLaunch {
  int rev_state = -1;  // state reversed in this launch, -1: none
  int rev_buf = -1;    // buffer holding it
  int nfwd = 0;
  AdjFwdOp fwd[kMaxChains];
};

struct AdjSchedule {
  std::vector<AdjLaunch> launches;  // the whole gradient: forward sweep, then reverse sweep
  int slots = 0;                    // snapshot slots used (peak number of stored states)
  int work = 0;                     // work buffers used
  long long forward_steps = 0;      // all forward steps, the first sweep included
  long long fused = 0;              // launches holding a reverse step and a forward step
  int levels = 0;                   // pipelined: number of levels (0 for binomial)
  std::vector<int> spacing;         // pipelined: level spacings K0 > K1 > ... > 1
};

// Binomial schedule for `states` = T + 1 states with `slots` snapshot slots (slot 0 = R_0).
// Returns false (with a message) if the arguments are invalid or the table is too large.
bool build_binomial(int states, int slots, AdjSchedule* out, std::string* why);

// Pipelined schedule with the given level spacings (K0 > K1 > ... > K_{L-1} = 1, each a
// multiple of the next; at most kMaxChains + 1 levels).
bool build_pipelined(int states, const std::vector<int>& spacing, AdjSchedule* out,
                     std::string* why);

// Pipelined schedule within a snapshot budget: the fewest levels (2 .. kMaxChains + 1) whose
// best spacing needs at most `slots` stored states.
bool build_pipelined_budget(int states, int slots, AdjSchedule* out, std::string* why);

// Replays a schedule on buffer contents (state indices) and checks that every reverse step
// and every forward step reads the state it expects, that the reverse steps run T, ..., 0,
// and that no launch writes a buffer it also reads.
bool check_schedule(int states, const AdjSchedule& s, std::string* why);

}  // namespace lbmhip
