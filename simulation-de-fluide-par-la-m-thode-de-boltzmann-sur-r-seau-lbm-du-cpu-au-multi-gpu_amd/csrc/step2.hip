// Two fused time steps per launch (temporal blocking), periodic domain with a uniform force:
// R_{t+2} from R_t with every population read once from HBM and written once, i.e. 72 B (fp32)
// per two cell updates instead of per one. Each cell update performs exactly the operations of
// the one-step kernel (the same pull, the same bounce-back rule, the same cell update: the
// reference's collide() for float and double, step_fp16_kernel's collision and rounding for
// FP16 storage), so the result is bit-identical to two single steps
// (src/v3_cuda_opt/solver_v3.cu:45-113 applied twice).
//
// A workgroup owns a tile of W = blockDim.x columns by H rows of the output and walks its rows
// downwards (in the +y direction). At row r it
//   1. computes step 1 (R_t -> R_{t+1}) at row r for its W columns, pulling from HBM (plane i
//      from row r - c_y,i of R_t: each plane row is read once per tile);
//   2. publishes the six populations its neighbour columns pull in step 2 (planes with c_x != 0)
//      in a four-row ring in LDS;
//   3. computes step 2 (R_{t+1} -> R_{t+2}) at row r - 1: planes with c_x = 0 come from the
//      thread's own registers (rows r - 2, r - 1, r), the others from the ring (columns x - 1,
//      x + 1), bounce-back from the thread's own R_{t+1} values of row r - 1; and stores it.
// Step 1 therefore runs on rows y0 - 1 .. y1 (one extra row on each side of the tile) and the
// two columns just outside the tile (x0 - 1 and x0 + W) are computed for 8 rows at a time
// (only the three populations the tile pulls from each of them are kept). Each ring row holds
// W + 2 columns: column 0 is the left halo column, 1 .. W the tile, W + 1 the right halo, so a
// lane reads its x - 1 and x + 1 neighbours at fixed offsets, the tile edges included. One
// barrier per row: with four ring slots a slot is rewritten two barriers after its last read.
// The recomputed rows and columns are the overhead of the scheme (rows y0 - 1 and y1 and their
// sources y0 - 2 and y1 + 1 are read by two tiles); step2_tile_rows() picks tiles of up to 64
// rows whose count fills the resident workgroup slots evenly. Solid cells are never written
// (both population buffers keep the values every upload gives them); their results go to
// discard slots past each plane's last cell.
#include <algorithm>
#include <type_traits>

#include "batch.hpp"
#include "half_cell.hpp"
#include "kernels.hpp"
#include "step_body.hpp"

namespace lbmhip {
namespace {

constexpr int kRingSlots = 4;
// Planes published in the ring and their index there: the ones neighbours pull (c_x != 0).
__host__ __device__ constexpr int ring_index(int plane) {
  return plane == 1 ? 0 : plane == 3 ? 1 : plane == 5 ? 2 : plane == 6 ? 3 : plane == 7 ? 4 : 5;
}
constexpr int kRingPlanes = 6;
// Planes kept of the halo columns: the left one (x0 - 1) gives the tile its c_x = +1 planes,
// the right one (x0 + W) its c_x = -1 planes.
constexpr int kLeftPlanes[3] = {1, 5, 8};
constexpr int kRightPlanes[3] = {3, 6, 7};
constexpr uint32_t kLinkBounce = 0x1feu;  // bits 1..8 of a link code

// Byte offsets of the source rows of step-1 row r (wrapped) and of the source columns of c.
__device__ __forceinline__ Rows rows_of(int r, int nx, int ny, unsigned B) {
  // (r == 0 ? ny : r) - 1 rather than r == 0 ? ny - 1 : r - 1: the compiler matches the latter
  // to a subtraction with borrow and selects it with vector instructions, uniform r or not
  const int rm = (r == 0 ? ny : r) - 1, rp = r == ny - 1 ? 0 : r + 1;
  return Rows{static_cast<unsigned>(rm * nx) * B, static_cast<unsigned>(r * nx) * B,
              static_cast<unsigned>(rp * nx) * B};
}
// The same for a row that is wave-uniform, stated so: where the compiler cannot prove an offset
// uniform it keeps it per lane, and a per-lane scalar offset costs a loop per load.
__device__ __forceinline__ Rows uniform_rows_of(int r, int nx, int ny, unsigned B) {
  const Rows a = rows_of(r, nx, ny, B);
  auto u = [](unsigned x) { return static_cast<unsigned>(__builtin_amdgcn_readfirstlane(x)); };
  return Rows{u(a.m), u(a.c), u(a.p)};
}
__device__ __forceinline__ Cols cols_of(int c, int nx, unsigned B) {
  return Cols{static_cast<unsigned>(c == 0 ? nx - 1 : c - 1) * B, static_cast<unsigned>(c) * B,
              static_cast<unsigned>(c == nx - 1 ? 0 : c + 1) * B};
}

// Segment flag at a wave-uniform byte index, by a scalar load of the aligned dword holding it
// (flag arrays are allocated in whole dwords): the branch on it is a scalar branch, and the
// wait for it does not wait for the wave's outstanding vector stores as a vector load's wait
// would.
__device__ __forceinline__ uint32_t uniform_flag(const uint8_t* flags, size_t index) {
  using ConstU32 = const __attribute__((address_space(4))) uint32_t*;
  const size_t at = static_cast<size_t>(__builtin_amdgcn_readfirstlane(static_cast<unsigned>(index)));
  const uintptr_t a = reinterpret_cast<uintptr_t>(flags) + at;
  const uint32_t w = *(ConstU32)(a & ~uintptr_t(3));
  return (w >> ((a & 3u) * 8u)) & 0xffu;
}

// A descriptor read with scalar loads: the caller's index is wave-uniform and the table is not
// written while kernels run.
template <class S>
__device__ __forceinline__ S load_uniform(const S* ptr) {
  static_assert(sizeof(S) % 4 == 0, "descriptor of whole dwords");
  using ConstU32 = const __attribute__((address_space(4))) uint32_t*;
  const ConstU32 src = (ConstU32)(ptr);
  uint32_t w[sizeof(S) / 4];
#pragma unroll
  for (size_t i = 0; i < sizeof(S) / 4; ++i) w[i] = src[i];
  return __builtin_bit_cast(S, w);
}

// One periodic domain with a uniform force: the plain solver's grid or one batched domain.
// S is the stored population type.
template <class S>
struct Step2View {
  const S* g;
  S* gn;
  const uint16_t* links;
  const uint8_t* segflags;
  size_t N;  // plane stride, >= nx * ny + kPlaneTail
  int nx, ny, nseg;
};

// Cell updates of the walk. Stored type S, computed type C, written word Word; step1() turns a
// cell's pulled values into its state one step later as the next step will read it, step2()
// into the words the store writes. `cell` is the cell's index y * nx + x; step1() returns a word
// that the walk hands to step2() of the same cell (FP16: the cell's hash for the step pair).
// What an open-x domain or a per-cell force field adds to a walk (float / double cells only):
// the inlet equilibrium and the force planes (fx at force, fy at force + N).
template <class T>
struct WalkEdges {
  const T* force;
  T eq_in[kQ];
};

// float / double: the reference's collision, bit-identical to step_kernel.
template <class T>
struct PlainCells {
  using S = T;
  using Ring = T;  // the values the walk carries (registers and LDS)
  using Word = T;
  // step 2's bounce-back reads the step-1 row above from the LDS ring (6 fewer values in
  // registers: fp32 117 -> 111 VGPRs, fp64 195 -> 183)
  static constexpr bool kRowInRegisters = false;
  // double: the next row's pulls are issued after step 1 (before the barrier) rather than
  // before it, so they are not in flight across step 1's collision: 8192^2 fp64 70.8k -> 71.3k
  // MLUPS, 8192 x 6144 67.5k -> 70.1k; float and FP16 measured no different either way
  static constexpr bool kPullLate = sizeof(T) == 8;
  // rows whose pulls are in flight ahead of the current one: double 2 (8192^2 +0.2-0.3 %, issued
  // late: early measured the same), float 1 (2 measured 0.3-0.4 % slower, strip included)
  static constexpr int kDepth = sizeof(T) == 8 ? 2 : 1;
  T omega, fp, fx, fy;
  __device__ __forceinline__ uint32_t step1(Ring (&v)[kQ], uint32_t) const {
    collide(v, fx, fy, omega, fp);
    return 0u;
  }
  __device__ __forceinline__ void step2(Ring (&v)[kQ], uint32_t, Word (&out)[kQ]) const {
    collide(v, fx, fy, omega, fp);
#pragma unroll
    for (int i = 0; i < kQ; ++i) out[i] = v[i];
  }
  // with the cell's own force (a per-cell field)
  __device__ __forceinline__ void step1(Ring (&v)[kQ], T cfx, T cfy) const { collide(v, cfx, cfy, omega, fp); }
  __device__ __forceinline__ void step2(Ring (&v)[kQ], T cfx, T cfy, Word (&out)[kQ]) const {
    collide(v, cfx, cfy, omega, fp);
#pragma unroll
    for (int i = 0; i < kQ; ++i) out[i] = v[i];
  }
  template <int kAux, class PS>
  __device__ __forceinline__ static void store(const PS& gn, int i, unsigned vo, unsigned so, Word w) {
    gn.template store<T, kAux>(i, vo, so, w);
  }
};

// FP16 storage: the collision and the stochastic rounding of step_fp16_kernel, for a launch
// that starts on an even step 2p: step 1 rounds with the cell's hash of pair p as is, step 2 with
// the same hash rotated (step_random), so the result is bit-identical to two one-step launches;
// step 2 reads the rounded step-1 values, as the second launch would.
// The walk carries binary16 numbers (the stored values and the rounded step-1 values), in
// registers and in the LDS ring; they are converted only inside the cell update, where every
// use is a (mixed-precision) fused multiply-add that reads the 16-bit value directly.
struct HalfCells {
  using S = uint16_t;
  using Ring = _Float16;
  using Word = unsigned short;
  // the FP16 walk keeps the whole step-1 row in registers (86 VGPRs with one row of pulls in
  // flight; reading it from the ring measured 109)
  static constexpr bool kRowInRegisters = true;
  static constexpr bool kPullLate = false;
  // two rows of pulls in flight (binary16: 128 B per wave and load): 8192^2 +0.6-1.3 %, 95 VGPRs,
  // still five waves per SIMD
  static constexpr int kDepth = 2;
  HalfCollide col;
  uint32_t seed;  // seed of the step pair
  __device__ __forceinline__ uint32_t step1(Ring (&v)[kQ], uint32_t cell) const {
    float f[kQ];
#pragma unroll
    for (int i = 0; i < kQ; ++i) f[i] = static_cast<float>(v[i]);
    const float mass = fast_collide(col, f);
    const uint32_t h = hash_cell(cell, seed);
    unsigned short o[kQ];
    round_cell(f, mass, h, col.unit, o);
#pragma unroll
    for (int i = 0; i < kQ; ++i) v[i] = __builtin_bit_cast(_Float16, o[i]);
    return h;
  }
  __device__ __forceinline__ void step2(Ring (&v)[kQ], uint32_t h, Word (&out)[kQ]) const {
    float f[kQ];
#pragma unroll
    for (int i = 0; i < kQ; ++i) f[i] = static_cast<float>(v[i]);
    const float mass = fast_collide(col, f);
    round_cell(f, mass, step_random(h, 16u), col.unit, out);
  }
  template <int, class PS>
  __device__ __forceinline__ static void store(const PS& gn, int i, unsigned vo, unsigned so, Word w) {
    gn.store_bits(i, vo, so, w);
  }
};

// Step 1 at a halo-column cell (c, r) (both wrapped) whose row differs between the lanes of a
// wave: the row offsets go into the per-lane offset (a per-lane scalar offset would make the
// compiler loop over its values). The pull with its bounce-back corrections and the collision,
// exactly as update_column.
template <class P, class PS, bool kOpen = false, bool kField = false>
__device__ __forceinline__ void halo_step(const Step2View<typename P::S>& p, const P& cells,
                                          const PS& g, int c, int r, typename P::Ring (&v)[kQ],
                                          const WalkEdges<typename P::S>* ex = nullptr) {
  using S = typename P::S;
  constexpr unsigned B = sizeof(S);
  const Rows rows = rows_of(r, p.nx, p.ny, B);
  const Cols cols = cols_of(c, p.nx, B);
#pragma unroll
  for (int i = 0; i < kQ; ++i) v[i] = g.template load_v<S>(i, cols.of(kCx[i]) + rows.of(kCy[i]), 0u);
  uint32_t link = 0;
  if (p.segflags[static_cast<size_t>(r) * p.nseg + (c >> 6)]) {
    link = p.links[static_cast<size_t>(r) * p.nx + c];
    if (!(link & kLinkSolid)) {
#pragma unroll
      for (int i = 1; i < kQ; ++i)
        if (link & (1u << i)) v[i] = g.template load_v<S>(kOpp[i], cols.c + rows.c, 0u);
      if constexpr (kOpen) {  // apply_open_edges with per-lane rows (all in the vector offset)
        if (link & kLinkInlet) {
          v[1] = ex->eq_in[1];
          v[5] = ex->eq_in[5];
          v[8] = ex->eq_in[8];
        }
        if (link & kLinkOutlet) {
          const unsigned left = cols.c - B;
#pragma unroll
          for (int i : {3, 6, 7})
            v[i] = (link & outlet_bb_bit(i)) ? g.template load_v<S>(kOpp[i], left + rows.c, 0u)
                                             : g.template load_v<S>(i, cols.c + rows.of(kCy[i]), 0u);
        }
      }
    }
  }
  if (!(link & kLinkSolid)) {
    if constexpr (kField) {
      const S cfx = Stored<S>::load(make_rsrc(ex->force), cols.c + rows.c, 0u);
      const S cfy = Stored<S>::load(make_rsrc(ex->force + p.N), cols.c + rows.c, 0u);
      cells.step1(v, cfx, cfy);
    } else {
      (void)cells.step1(v, static_cast<uint32_t>(r) * static_cast<uint32_t>(p.nx) + static_cast<uint32_t>(c));
    }
  }
}

// Halo-column cells are computed for kHaloChunk step-1 rows at a time (the first chunk also
// covers the two leading rows), so the LDS a workgroup needs does not grow with its tile height.
// A multiple of 4 (the main loop's unrolling). The shorter the chunk, the closer to the time
// the neighbour tile reads its lines the halo columns read them: 8192^2 FP16 229.1k MLUPS with
// 8 rows against 224.8k with 16 and 224.1k with 4 (fp32 148.6k, 148.5k, 148.3k; 32 rows was
// 0.6 % slower than 16); same bits for any chunk.
constexpr int kHaloChunk = 8;
constexpr int kHaloRows = kHaloChunk + 2;

// The order of a walk over nr step-1 rows: the first halo chunk (kHaloRows rows), start()
// (no barrier), rows 0 and 1, then four rows per trip with constant ring slots (2, 3, 0, 1) and a remainder, a halo chunk of
// kHaloChunk rows before the row that has used up the previous one (trip starts are 2 mod 4,
// and so are the chunk starts 18, 34, ...). In a walk every halo_chunk(k0, cnt) and
// row(k, hk, slot, two) ends or contains exactly one barrier, so step2_barriers(nr) counts a
// walk's barriers from this same schedule.
template <class Halo, class Start, class Row>
__device__ __forceinline__ void step2_schedule(const int nr, Halo&& halo_chunk, Start&& start,
                                               Row&& row) {
  halo_chunk(0, min(kHaloRows, nr));
  start();
  row(0, 0, 0, false);
  row(1, 1, 1, false);
  // Four rows per trip: constant ring slots (LDS offsets), and the copies between the rows'
  // register sets disappear (unrolling by 1 or 2, or capping the registers at 80 for a fourth
  // fp32 workgroup per CU, measured 3-5 % slower). A chunk's entries are dead once its last
  // row has passed its barrier.
  int k = 2, hk = 2, hcap = kHaloRows;
  for (; k + 4 <= nr; k += 4, hk += 4) {
    if (hk == hcap) {
      halo_chunk(k, min(kHaloChunk, nr - k));
      hk = 0;
      hcap = kHaloChunk;
    }
    row(k, hk, 2, true);
    row(k + 1, hk + 1, 3, true);
    row(k + 2, hk + 2, 0, true);
    row(k + 3, hk + 3, 1, true);
  }
  for (; k < nr; ++k, ++hk) {
    if (hk == hcap) {
      halo_chunk(k, min(kHaloChunk, nr - k));
      hk = 0;
      hcap = kHaloChunk;
    }
    row(k, hk, k & (kRingSlots - 1), true);
  }
}

// Barriers of a walk over nr step-1 rows (wave-uniform).
__device__ __forceinline__ int step2_barriers(const int nr) {
  int n = 0;
  step2_schedule(nr, [&](int, int) { ++n; }, [] {}, [&](int, int, int, bool) { ++n; });
  return n;
}

// The walk of one tile by W threads (thread t = 0 .. W - 1, W a multiple of 64, whole waves):
// columns x0 .. x0 + W - 1 (those inside the grid), output rows y0 .. y0 + H - 1 (those below
// row_end), with the LDS region at smem (step2_region_elems(W) values). Ends with every LDS
// read done before the next walk's first LDS write crosses a barrier (its halo chunk ends with
// one). Its barriers: step2_barriers(min(y0 + H, row_end) - y0 + 2).
// A whole periodic domain has row_end = ny; a strip's owned rows [row0, row_end) with two ghost
// rows on each side (row0 >= 2, row_end <= ny - 2) read rows row0 - 2 .. row_end + 1 only, so
// no row wraps and the ghost rows are plain inputs.
// kOpen: an open-x domain (inlet column 0, outlet column nx - 1: the rules of apply_links in
// both steps, the second one's from the ring); kField: a per-cell force field (each row's force
// loaded with its pulls, kept for its step 2). Both take ex and float / double cells.
template <class P, class PS, int kStoreAux, bool kHalo = false, bool kWhole = false,
          bool kAhead = false, bool kOpen = false, bool kField = false>
__device__ __forceinline__ void step2_walk(const Step2View<typename P::S>& p, const P& cells,
                                           const int x0, const int y0, const int H,
                                           const int row_end, unsigned char* smem, const int W,
                                           const int t,
                                           const HaloOut<typename P::S>* hout = nullptr,
                                           const WalkEdges<typename P::S>* ex = nullptr) {
  static_assert(!(kOpen || kField) || (!P::kRowInRegisters && !kAhead && !kWhole),
                "open edges and force fields: plain cells, bounce-back data in its row");
  using T = typename P::Ring;  // the values the walk carries, in registers and in LDS
  using S = typename P::S;
  const int WP = W + 2;
  T* ring = reinterpret_cast<T*>(smem);                  // [kRingSlots][kRingPlanes][W + 2]
  T* halo = ring + kRingSlots * kRingPlanes * WP;        // [2][kHaloRows][3]
  const int nx = p.nx, ny = p.ny;
  const int wv = min(W, nx - x0);  // valid columns of this tile
  const int y1 = min(y0 + H, row_end);
  const int nr = y1 - y0 + 2;                   // step-1 rows y0 - 1 .. y1
  PS g, gn;
  g.init(p.g, p.N);
  gn.init(p.gn, p.N);
  auto wrap_y = [&](int r) { return r < 0 ? r + ny : (r >= ny ? r - ny : r); };

  // Step 1 on the two halo columns for step-1 rows k0 .. k0 + cnt - 1 (cnt <= kHaloRows), into
  // halo[side][k - k0]; then a barrier. A loop with a wave-uniform trip count: values the
  // compiler shares with the main loop stay scalar.
  const int cl = x0 == 0 ? nx - 1 : x0 - 1, cr = x0 + wv == nx ? 0 : x0 + wv;
  // kWhole (x0 = 0 and W >= nx): the tile is as wide as its grid and is its own halo. The left
  // neighbour of column 0 is column nx - 1 and the right one of column nx - 1 is column 0, both
  // computed by the walk itself (lanes wv - 1 and 0 copy them into the ring's halo columns), so
  // its halo chunks only keep their barrier. (The values of a solid cell differ, the walk
  // colliding it and halo_step not, but no step reads them: its fluid neighbours bounce back.)
  constexpr bool whole = kWhole;
  auto halo_chunk = [&](const int k0, const int cnt) __attribute__((always_inline)) {
    for (int j0 = 0; j0 < (whole ? 0 : 2 * cnt); j0 += W) {
      const int j = j0 + t;
      if (j < 2 * cnt) {
        const int side = j >= cnt ? 1 : 0, kk = j - side * cnt;
        T v[kQ];
        halo_step<P, PS, kOpen, kField>(p, cells, g, side ? cr : cl, wrap_y(y0 - 1 + k0 + kk), v, ex);
        T* h = halo + (side * kHaloRows + kk) * 3;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          // both candidates as register values (an empty asm): a select of two elements of v
          // would otherwise become v[index] with the index selected, a compare chain per plane
          T a = v[kLeftPlanes[q]], b = v[kRightPlanes[q]];
          asm volatile("" : "+v"(a), "+v"(b));
          h[q] = side ? b : a;
        }
      }
    }
    __syncthreads();
  };

  // Idle lanes (past the tile's last column) shadow the last column in everything, their ring
  // column included: they compute its values and store them to its address, the same bits.
  const int tt = t < wv ? t : wv - 1;
  const int c = x0 + tt;
  // the wave's 64-column segment of the flag array: a tile starts on a multiple of 64 columns,
  // so every lane of a wave (idle ones included) lies in lane 0's segment; taken from lane 0,
  // the flag index is scalar arithmetic
  const int seg = __builtin_amdgcn_readfirstlane(c) >> 6;
  constexpr unsigned B = sizeof(S);
  const unsigned vo = static_cast<unsigned>(c) * B;
  // A solid cell is not written: its step-2 result goes to a discard slot past the plane's
  // last cell (a per-lane offset instead of a branch around the stores, see below).
  const unsigned tail = static_cast<unsigned>(nx * ny) * B + static_cast<unsigned>(t & 63) * B;
  const Cols cols = cols_of(c, nx, B);
  // R_{t+1} at (c, r - 1): all planes (kRowInRegisters), or those with c_x = 0 (the others are
  // in the ring slot of row r - 1, where the bounce-back of rows next to a wall reads them)
  T prev[kQ];
#pragma unroll
  for (int i = 0; i < kQ; ++i) prev[i] = T(0);
  T prev2 = T(0);       // plane 2 of R_{t+1} at (c, r - 2)
  uint32_t prev_link = kLinkSolid;
  // wave-uniform: the segment flag of row r - 1 (0: every link code of the wave's cells there
  // is 0, so its step 2 needs no bounce-back, open-edge or solid-cell handling)
  uint32_t prev_flag = 1;
  uint32_t prev_carry = 0;  // what step 1 at (c, r - 1) handed to its step 2
  // pulls in flight: of the next row (kDepth 1), or of the next two rows (kDepth 2: more bytes
  // in flight per wave, at the cost of nine registers; the batched walk's bounce-back data and
  // a force field's rows are loaded one row ahead, so those walks keep one)
  constexpr int kDepth = (kAhead || kField) ? 1 : P::kDepth;
  T cur[kQ];            // raw pulls of the current step-1 row
  T nx1[kQ];            // kDepth 2: raw pulls of the next row
  auto pull = [&](const Rows& rw, T (&out)[kQ]) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) out[i] = g.template load_v<S>(i, cols.of(kCx[i]), rw.of(kCy[i]));
  };
  // kAhead: the bounce-back data of a step-1 row (its segment flag, link codes and own values)
  // is loaded with the pulls, one row before its use, so that a row next to a wall or an
  // obstacle does not wait for its loads; otherwise it is loaded and waited for in its row.
  uint32_t bflag = 0, blink = 0;
  T bown[kQ];
  if constexpr (kAhead) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) bown[i] = T(0);
  }
  auto bounce_data = [&](const int r, const Rows& rw) __attribute__((always_inline)) {
    bflag = uniform_flag(p.segflags, static_cast<size_t>(r) * p.nseg + seg);
    if (bflag) {
      blink = p.links[static_cast<size_t>(r) * nx + c];
#pragma unroll
      for (int i = 1; i < kQ; ++i) bown[i] = g.template load_v<S>(kOpp[i], cols.c, rw.c);
    }
  };
  // kField: the force of step-1 row r (f1), of the row above it (f2, step 2's), of the next row
  const Rsrc fxr = make_rsrc(kField ? static_cast<const void*>(ex->force) : nullptr);
  const Rsrc fyr = make_rsrc(kField ? static_cast<const void*>(ex->force + p.N) : nullptr);
  S f1x = S(0), f1y = S(0), f2x = S(0), f2y = S(0);
  // after the first halo chunk: the pulls of step-1 row 0
  auto start = [&] {
    const int r0 = wrap_y(y0 - 1);
    const Rows rows0 = uniform_rows_of(r0, nx, ny, B);
    pull(rows0, cur);
    if constexpr (kDepth == 2) pull(uniform_rows_of(wrap_y(y0), nx, ny, B), nx1);
    if constexpr (kAhead) bounce_data(r0, rows0);
    if constexpr (kField) {
      f1x = Stored<S>::load(fxr, cols.c, rows0.c);
      f1y = Stored<S>::load(fyr, cols.c, rows0.c);
    }
  };

  // Step-1 row k (ring slot k & 3) and, when `two`, step 2 at row k - 1. The pulls of row
  // k + kDepth are issued first (P::kPullLate: after step 1, before the barrier) and nothing
  // between them and their use branches around a vector memory operation (the bounce-back loads
  // complete inside their own branch, solid cells store to the discard slots): one wait counter
  // covers loads and stores, and a branch around either would make the wait for the pulls a
  // wait for every store in flight.
  auto row = [&](const int k, const int hk, const int slot, const bool two)
                 __attribute__((always_inline)) {
    const int r = wrap_y(y0 - 1 + k);
    const Rows rows = uniform_rows_of(r, nx, ny, B);
    T nxt[kQ];  // the last rows re-read their own rows (cached) instead of branching
    const int rn = k + 1 < nr ? wrap_y(y0 + k) : r;
    const Rows rows_n = uniform_rows_of(rn, nx, ny, B);
    // the row whose pulls this row issues
    const Rows rows_p = kDepth == 1 ? rows_n
                                    : uniform_rows_of(k + 2 < nr ? wrap_y(y0 + k + 1) : r, nx, ny, B);
    S fnx = S(0), fny = S(0);
    auto issue_next = [&]() __attribute__((always_inline)) {
      pull(rows_p, nxt);
      if constexpr (kField) {
        fnx = Stored<S>::load(fxr, cols.c, rows_n.c);
        fny = Stored<S>::load(fyr, cols.c, rows_n.c);
      }
    };
    if constexpr (!P::kPullLate) issue_next();
    // step 1 at (c, r), exactly as update_column: bounce-back corrections, then the collision
    // (of a solid cell too: its values are never read, its fluid neighbours bounce back)
    T v[kQ];
#pragma unroll
    for (int i = 0; i < kQ; ++i) v[i] = cur[i];
    uint32_t link = 0, flag;
    if constexpr (kAhead) {
      flag = bflag;
      link = bflag ? blink : 0u;
      if (bflag) {
#pragma unroll
        for (int i = 1; i < kQ; ++i) v[i] = (link & (1u << i)) ? bown[i] : v[i];
      }
      // the next row's (its loads complete by the next row's wait for its pulls)
      if (!P::kPullLate && k + 1 < nr) bounce_data(rn, rows_n);
    } else if ((flag = uniform_flag(p.segflags, static_cast<size_t>(r) * p.nseg + seg))) {
      link = p.links[static_cast<size_t>(r) * nx + c];
      T own[kQ];
#pragma unroll
      for (int i = 1; i < kQ; ++i) own[i] = g.template load_v<S>(kOpp[i], cols.c, rows.c);
#pragma unroll
      for (int i = 1; i < kQ; ++i) v[i] = (link & (1u << i)) ? own[i] : v[i];
      if constexpr (kOpen) apply_open_edges<S>(link, g, rows, cols, ex->eq_in, v);
    }
    uint32_t carry = 0;
    if constexpr (kField)
      cells.step1(v, f1x, f1y);
    else
      carry = cells.step1(v, static_cast<uint32_t>(r) * static_cast<uint32_t>(nx) + static_cast<uint32_t>(c));
    T* sl = ring + slot * kRingPlanes * WP;
#pragma unroll
    for (int i : {1, 3, 5, 6, 7, 8}) sl[ring_index(i) * WP + tt + 1] = v[i];
    if (t == 0) {
      if (whole) {
#pragma unroll
        for (int q = 0; q < 3; ++q) sl[ring_index(kRightPlanes[q]) * WP + wv + 1] = v[kRightPlanes[q]];
      } else {
        const T* h = halo + hk * 3;
#pragma unroll
        for (int q = 0; q < 3; ++q) sl[ring_index(kLeftPlanes[q]) * WP] = h[q];
      }
    }
    if (t == wv - 1) {
      if (whole) {
#pragma unroll
        for (int q = 0; q < 3; ++q) sl[ring_index(kLeftPlanes[q]) * WP] = v[kLeftPlanes[q]];
      } else {
        const T* h = halo + (kHaloRows + hk) * 3;
#pragma unroll
        for (int q = 0; q < 3; ++q) sl[ring_index(kRightPlanes[q]) * WP + wv + 1] = h[q];
      }
    }
    if constexpr (P::kPullLate) {
      issue_next();
      if constexpr (kAhead) {
        if (k + 1 < nr) bounce_data(rn, rows_n);
      }
    }
    __syncthreads();
    if (two) {
      // step 2 at row q = r - 1 (an output row of the tile); ring column tt is x - 1, tt + 2
      // is x + 1
      const T* s0 = sl;                                                  // row q + 1
      const T* s1 = ring + ((slot + 3) & (kRingSlots - 1)) * kRingPlanes * WP;  // row q
      const T* s2 = ring + ((slot + 2) & (kRingSlots - 1)) * kRingPlanes * WP;  // row q - 1
      T w[kQ];
      w[0] = prev[0];
      w[1] = s1[ring_index(1) * WP + tt];
      w[2] = prev2;
      w[3] = s1[ring_index(3) * WP + tt + 2];
      w[4] = v[4];
      w[5] = s2[ring_index(5) * WP + tt];
      w[6] = s2[ring_index(6) * WP + tt + 2];
      w[7] = s0[ring_index(7) * WP + tt + 2];
      w[8] = s0[ring_index(8) * WP + tt];
      const int q = y0 - 2 + k;
      const unsigned so = static_cast<unsigned>(
          __builtin_amdgcn_readfirstlane(static_cast<unsigned>(q * nx) * B));
      // the store offset of solid cells (their results go to the discard slots): taken with the
      // other link rules, or after the collision where bounce-back data is held a row ahead
      // (kAhead: computed before the collision it cost the batched walk 10 VGPRs, a wave per
      // SIMD and 9 % of its rate)
      unsigned vs = vo;
      if (prev_flag) {  // wave-uniform: a wall, an obstacle or an open edge in the segment
        const uint32_t bb = prev_link & kLinkBounce;
        if (__builtin_amdgcn_ballot_w64(bb != 0)) {  // wave-uniform
          // R_{t+1} at (c, q)
          T pq[kQ];
#pragma unroll
          for (int i = 0; i < kQ; ++i) pq[i] = prev[i];
          if constexpr (!P::kRowInRegisters) {
#pragma unroll
            for (int i : {1, 3, 5, 6, 7, 8}) pq[i] = s1[ring_index(i) * WP + tt + 1];
          }
#pragma unroll
          for (int i = 1; i < kQ; ++i)
            if (bb & (1u << i)) w[i] = pq[kOpp[i]];
        }
        if constexpr (kOpen) {
          // apply_open_edges on the step-1 values: the inlet equilibrium, and at the outlet
          // what x - 1 receives, R_{t+1}(i, x, q - c_y) (ring rows q, q - 1, q + 1 at column
          // x) or R_{t+1}(opp(i), x - 1, q) (ring row q at column x - 1)
          if (prev_link & kLinkInlet) {
            w[1] = ex->eq_in[1];
            w[5] = ex->eq_in[5];
            w[8] = ex->eq_in[8];
          }
          if (prev_link & kLinkOutlet) {
            w[3] = (prev_link & outlet_bb_bit(3)) ? s1[ring_index(kOpp[3]) * WP + tt]
                                                  : s1[ring_index(3) * WP + tt + 1];
            w[6] = (prev_link & outlet_bb_bit(6)) ? s1[ring_index(kOpp[6]) * WP + tt]
                                                  : s2[ring_index(6) * WP + tt + 1];
            w[7] = (prev_link & outlet_bb_bit(7)) ? s1[ring_index(kOpp[7]) * WP + tt]
                                                  : s0[ring_index(7) * WP + tt + 1];
          }
        }
        if constexpr (!kAhead) vs = (prev_link & kLinkSolid) ? tail - so : vo;
      }
      typename P::Word out[kQ];
      if constexpr (kField)
        cells.step2(w, f2x, f2y, out);
      else
        cells.step2(w, prev_carry, out);
      if constexpr (kAhead) {
        if (prev_flag) vs = (prev_link & kLinkSolid) ? tail - so : vo;
      }
#pragma unroll
      for (int i = 0; i < kQ; ++i) P::template store<kStoreAux>(gn, i, vs, so, out[i]);
      if constexpr (kHalo) {
        // the halo entries of the sender's two outermost owned rows on each side (q is
        // wave-uniform: scalar branches, taken on four rows of a launch)
        const HaloOut<S>& h = *hout;
        if (q >= h.hi - 1 || q <= h.lo + 1) {
          const bool solid = (prev_link & kLinkSolid) != 0;
          const unsigned lane = static_cast<unsigned>(t & 63) * B;
          auto put = [&](S* base, unsigned stride, unsigned tail, int row, int plane) {
            const unsigned so_h = static_cast<unsigned>(row * nx) * B;
            const unsigned vo_h = solid ? tail + lane - so_h : vo;
            Stored<S>::template store<kAuxNone>(make_rsrc(base), vo_h,
                                                so_h + static_cast<unsigned>(plane) * stride,
                                                out[plane]);
          };
          if (q == h.hi) {  // all planes -> ghost row 1 of the strip above
#pragma unroll
            for (int i = 0; i < kQ; ++i) put(h.up, h.up_stride, h.up_tail, 1, i);
          }
          if (q == h.hi - 1) {  // c_y = +1 planes -> its ghost row 0
#pragma unroll
            for (int i : {2, 5, 6}) put(h.up, h.up_stride, h.up_tail, 0, i);
          }
          if (q == h.lo) {  // all planes -> the first ghost row above of the strip below
#pragma unroll
            for (int i = 0; i < kQ; ++i) put(h.down, h.down_stride, h.down_tail, h.down_row, i);
          }
          if (q == h.lo + 1) {  // c_y = -1 planes -> its outer ghost row
#pragma unroll
            for (int i : {4, 7, 8}) put(h.down, h.down_stride, h.down_tail, h.down_row + 1, i);
          }
        }
      }
    }
    prev2 = prev[2];
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      if (P::kRowInRegisters || kCx[i] == 0) prev[i] = v[i];
      if constexpr (kDepth == 1) {
        cur[i] = nxt[i];
      } else {
        cur[i] = nx1[i];
        nx1[i] = nxt[i];
      }
    }
    prev_link = link;
    prev_flag = flag;
    prev_carry = carry;
    if constexpr (kField) {
      f2x = f1x;
      f2y = f1y;
      f1x = fnx;
      f1y = fny;
    }
  };

  step2_schedule(nr, halo_chunk, start, row);
}

template <class T, class PS, int kStoreAux, bool kHalo = false, bool kOpen = false,
          bool kField = false>
__global__ void __launch_bounds__(512) step2_kernel(const StepParams<T> p, const HaloOut<T> h) {
  extern __shared__ __align__(16) unsigned char smem[];
  const unsigned tile = tile_in_order(p.xcd_bands, blockIdx.x, gridDim.x, p.tiles_x);
  const int tx = static_cast<int>(tile % static_cast<unsigned>(p.tiles_x));
  const int ty = static_cast<int>(tile / static_cast<unsigned>(p.tiles_x));
  const Step2View<T> d{p.g, p.gn, p.links, p.segflags, p.N, p.nx, p.ny, p.nseg};
  const PlainCells<T> cells{p.omega, p.fp, p.fx, p.fy};
  WalkEdges<T> ex;
  if constexpr (kOpen || kField) {
    ex.force = p.force;
#pragma unroll
    for (int i = 0; i < kQ; ++i) ex.eq_in[i] = p.eq_in[i];
  }
  step2_walk<PlainCells<T>, PS, kStoreAux, kHalo, false, false, kOpen, kField>(
      d, cells, tx * static_cast<int>(blockDim.x), p.row0 + ty * p.rows_per_tile, p.rows_per_tile,
      p.row0 + p.nrows, smem, static_cast<int>(blockDim.x), static_cast<int>(threadIdx.x), &h, &ex);
}

// FP16 storage, two steps from an even step: p.seed is the seed of the step pair.
template <class PS>
__global__ void __launch_bounds__(512) step2_fp16_kernel(const StepParamsH p) {
  extern __shared__ __align__(16) unsigned char smem[];
  const unsigned tile = tile_in_order(p.xcd_bands, blockIdx.x, gridDim.x,
                                      static_cast<unsigned>(p.tiles_x));
  const int tx = static_cast<int>(tile % static_cast<unsigned>(p.tiles_x));
  const int ty = static_cast<int>(tile / static_cast<unsigned>(p.tiles_x));
  const Step2View<uint16_t> d{p.g, p.gn, p.links, p.segflags, p.N, p.nx, p.ny, p.nseg};
  const HalfCells cells{p.col, p.seed};
  step2_walk<HalfCells, PS, kAuxNone>(d, cells, tx * static_cast<int>(blockDim.x),
                                      ty * p.rows_per_tile, p.rows_per_tile, p.ny, smem,
                                      static_cast<int>(blockDim.x), static_cast<int>(threadIdx.x));
}

// Batched domains: a grid of G workgroups walks the packed two-step work items, workgroup b
// taking items b, b + G, b + 2G, ... Each wave reads its own entry of the item (scalar loads):
// the waves of one sub-tile walk it together as a workgroup of `width` threads would, with
// their own LDS region; then every wave pads its barriers to those of the item's tallest
// sub-tile (a wave without a sub-tile only pads), so that all waves of the workgroup meet the
// same number of barriers, and one more barrier ends the item before the next one's sub-tiles
// reuse the LDS.
template <class T, class PS, int kStoreAux>
__global__ void __launch_bounds__(512) batch_step2_kernel(const BatchParams<T> p) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nw = static_cast<int>(blockDim.x) >> 6;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
  for (int it = static_cast<int>(blockIdx.x); it < p.ntiles; it += static_cast<int>(gridDim.x)) {
    const BatchWave w = load_uniform(p.waves + static_cast<size_t>(it) * nw + wave);
    const int total = step2_barriers(w.nr_max);
    int done = 0;
    if (w.domain >= 0) {
      const BatchDomain<T> dm = load_uniform(p.domains + w.domain);
      const Step2View<T> d{p.g + dm.offset, p.gn + dm.offset, p.links + dm.cell_offset,
                           p.segflags + dm.seg_offset, dm.N, dm.nx, dm.ny, dm.nseg};
      const PlainCells<T> cells{dm.omega, dm.fp, dm.fx, dm.fy};
      unsigned char* region = smem + static_cast<size_t>(w.lds) * sizeof(T);
      const int lt = static_cast<int>(threadIdx.x) - w.lane0;
      // bounce-back data one row ahead: walls and obstacles are common in small domains
      if (w.x0 == 0 && w.width >= dm.nx)  // a whole domain's width: its own halo
        step2_walk<PlainCells<T>, PS, kStoreAux, false, true, true>(d, cells, 0, w.y0, w.rows,
                                                                    dm.ny, region, w.width, lt);
      else
        step2_walk<PlainCells<T>, PS, kStoreAux, false, false, true>(d, cells, w.x0, w.y0, w.rows,
                                                                     dm.ny, region, w.width, lt);
      done = step2_barriers(w.rows + 2);
    }
    for (; done < total; ++done) __syncthreads();
    __syncthreads();
  }
}

template <class T, class PS, bool kOpen, bool kField>
void launch_edges(const StepParams<T>& p, unsigned grid, int block_x, size_t lds, hipStream_t stream) {
  const HaloOut<T> none{};
  if (p.store_nt)
    step2_kernel<T, PS, kAuxNt, false, kOpen, kField><<<grid, block_x, lds, stream>>>(p, none);
  else
    step2_kernel<T, PS, kAuxNone, false, kOpen, kField><<<grid, block_x, lds, stream>>>(p, none);
}

template <class T, class PS>
void launch_planes(const StepParams<T>& p, bool open_x, unsigned grid, int block_x, size_t lds,
                   hipStream_t stream) {
  const bool field = p.force != nullptr;
  if (open_x && field)
    launch_edges<T, PS, true, true>(p, grid, block_x, lds, stream);
  else if (open_x)
    launch_edges<T, PS, true, false>(p, grid, block_x, lds, stream);
  else if (field)
    launch_edges<T, PS, false, true>(p, grid, block_x, lds, stream);
  else
    launch_edges<T, PS, false, false>(p, grid, block_x, lds, stream);
}

// The checks of launch_step2 (one function for both entry points); returns the grid size, or 0
// for an invalid configuration.
template <class T>
unsigned step2_grid(const StepParams<T>& p, int block_x, bool edges = false) {
  // a whole periodic-y domain, or the owned rows of a strip with two ghost rows on each side
  // (uniform force, periodic x); tile rows >= 1; a plane addressable with 32 bits and holding
  // the discard slots. `edges`: open x or a force field allowed (whole domains only).
  const bool whole = p.wrap_y && p.row0 == 0 && p.nrows == p.ny;
  const bool strip = !p.wrap_y && p.row0 >= 2 && p.nrows >= 1 && p.row0 + p.nrows <= p.ny - 2;
  if (!(whole || strip) || (p.force && !(edges && whole)) || p.rows_per_tile < 1 ||
      block_x < 64 || block_x > 512 || block_x % 64 ||
      p.tiles_x != (p.nx + block_x - 1) / block_x ||
      static_cast<uint64_t>(p.N) * sizeof(T) >= (uint64_t{1} << 32) ||
      p.N < static_cast<size_t>(p.nx) * p.ny + kPlaneTail)
    return 0;
  if (step2_lds_bytes(block_x, sizeof(T)) > 64 * 1024) return 0;
  const unsigned tiles_y = static_cast<unsigned>((p.nrows + p.rows_per_tile - 1) / p.rows_per_tile);
  if (p.xcd_bands == 2 && tiles_y % 8) return 0;
  return static_cast<unsigned>(p.tiles_x) * tiles_y;
}

}  // namespace

size_t step2_region_elems(int width) {
  return static_cast<size_t>(kRingSlots) * kRingPlanes * (width + 2) + 2 * kHaloRows * 3;
}

size_t step2_lds_bytes(int block_x, size_t elem) { return step2_region_elems(block_x) * elem; }

size_t batch_step2_lds_bytes(int block_x, size_t elem) {
  // the most sub-tiles a workgroup can hold: one per wave
  return (static_cast<size_t>(kRingSlots) * kRingPlanes * block_x +
          static_cast<size_t>(block_x / 64) * (2 * kRingSlots * kRingPlanes + 2 * kHaloRows * 3)) *
         elem;
}

template <class T>
hipError_t launch_step2(const StepParams<T>& p, int block_x, hipStream_t stream, bool open_x) {
  const unsigned grid = step2_grid(p, block_x, true);
  // open x: a whole domain at least 3 columns wide (inlet, outlet and a column between)
  if (!grid || (open_x && (!p.wrap_y || p.nx < 3))) return hipErrorInvalidValue;
  const size_t lds = step2_lds_bytes(block_x, sizeof(T));
  const bool one = p.one_rsrc && static_cast<uint64_t>(kQ) * p.N * sizeof(T) < (uint64_t{1} << 32);
  if (one)
    launch_planes<T, PlanesOne>(p, open_x, grid, block_x, lds, stream);
  else
    launch_planes<T, Planes>(p, open_x, grid, block_x, lds, stream);
  return hipGetLastError();
}
template hipError_t launch_step2<float>(const StepParams<float>&, int, hipStream_t, bool);
template hipError_t launch_step2<double>(const StepParams<double>&, int, hipStream_t, bool);

template <class T>
hipError_t launch_step2_halo(const StepParams<T>& p, const HaloOut<T>& h, int block_x,
                             hipStream_t stream) {
  const unsigned grid = step2_grid(p, block_x);
  // a strip (no y wrap) whose owned rows include h.lo .. h.hi's neighbours as the kernel sees
  // them; every receiver's planes in one buffer resource
  if (!grid || p.wrap_y || !h.up || !h.down || h.lo < 2 || h.hi < h.lo + 1 || h.hi > p.ny - 3 ||
      h.down_row < 3 || h.up_tail < 2 * static_cast<unsigned>(p.nx) * sizeof(T) ||
      h.down_tail < static_cast<unsigned>(h.down_row + 2) * static_cast<unsigned>(p.nx) * sizeof(T) ||
      static_cast<uint64_t>(kQ) * h.up_stride >= (uint64_t{1} << 32) ||
      static_cast<uint64_t>(kQ) * h.down_stride >= (uint64_t{1} << 32))
    return hipErrorInvalidValue;
  const size_t lds = step2_lds_bytes(block_x, sizeof(T));
  const bool one = p.one_rsrc && static_cast<uint64_t>(kQ) * p.N * sizeof(T) < (uint64_t{1} << 32);
  if (one)
    step2_kernel<T, PlanesOne, kAuxNt, true><<<grid, block_x, lds, stream>>>(p, h);
  else
    step2_kernel<T, Planes, kAuxNt, true><<<grid, block_x, lds, stream>>>(p, h);
  return hipGetLastError();
}
template hipError_t launch_step2_halo<float>(const StepParams<float>&, const HaloOut<float>&, int,
                                             hipStream_t);
template hipError_t launch_step2_halo<double>(const StepParams<double>&, const HaloOut<double>&,
                                              int, hipStream_t);

template <class T>
hipError_t launch_batch_step2(const BatchParams<T>& p, int block_x, int workgroups,
                              hipStream_t stream) {
  // every domain periodic with a uniform force and the plane tail (the batch layout has both)
  if (p.ntiles <= 0) return hipSuccess;
  if (!p.waves || block_x < 64 || block_x > 512 || block_x % 64 || workgroups < 0)
    return hipErrorInvalidValue;
  const size_t lds = batch_step2_lds_bytes(block_x, sizeof(T));
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  const unsigned grid =
      static_cast<unsigned>(workgroups > 0 ? std::min(workgroups, p.ntiles) : p.ntiles);
  if (p.one_rsrc) {
    if (p.store_nt)
      batch_step2_kernel<T, PlanesOne, kAuxNt><<<grid, block_x, lds, stream>>>(p);
    else
      batch_step2_kernel<T, PlanesOne, kAuxNone><<<grid, block_x, lds, stream>>>(p);
  } else {
    if (p.store_nt)
      batch_step2_kernel<T, Planes, kAuxNt><<<grid, block_x, lds, stream>>>(p);
    else
      batch_step2_kernel<T, Planes, kAuxNone><<<grid, block_x, lds, stream>>>(p);
  }
  return hipGetLastError();
}
template hipError_t launch_batch_step2<float>(const BatchParams<float>&, int, int, hipStream_t);
template hipError_t launch_batch_step2<double>(const BatchParams<double>&, int, int, hipStream_t);

template <class T>
int batch_step2_resident(int block_x, bool store_nt, bool one_rsrc) {
  const size_t lds = batch_step2_lds_bytes(block_x, sizeof(T));
  int n = 0;
  hipError_t e;
  if (one_rsrc)
    e = store_nt ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_step2_kernel<T, PlanesOne, kAuxNt>, block_x, lds)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_step2_kernel<T, PlanesOne, kAuxNone>, block_x, lds);
  else
    e = store_nt ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_step2_kernel<T, Planes, kAuxNt>, block_x, lds)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_step2_kernel<T, Planes, kAuxNone>, block_x, lds);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}
template int batch_step2_resident<float>(int, bool, bool);
template int batch_step2_resident<double>(int, bool, bool);

// Tile height for a two-step kernel instance: `kernel` with block_x threads and `lds` bytes.
template <class K>
int tile_rows_for(K kernel, int nx, int ny, int block_x, size_t lds, int min_waves = 1) {
  // Workgroups of this instance resident at once on the current device.
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block_x, lds) != hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      per_cu < 1 || cus < 1) {
    (void)hipGetLastError();
    return kHaloChunk;
  }
  // The tallest candidate whose tiles fill the resident slots evenly (the last wave of
  // workgroups at least 97 % full), else the best-filling one. 8192^2 fp32, 512 columns, two
  // workgroups per CU: 64 rows (2048 tiles, four full waves) 146.3k MLUPS; 72 rows (1824 tiles,
  // last wave 56 % full) 125.8k, 56 rows 128.3k; 32 rows (eight waves, more recomputed rows)
  // 141.9k; 128 and 256 rows (two waves, one) 134.9k and 132.9k. With three waves or more a
  // partly filled last wave costs little (workgroups finish at different times): FP16 storage,
  // 256 columns, five workgroups per CU: 64 rows (3.2 waves) 210.6k against 205.8k for the
  // 16 rows that fill the slots evenly. min_waves: candidates with fewer full waves are passed
  // over unless none has that many (a strip's 8192 x 1024 owned rows: 16 rows, two waves,
  // 136.2k MLUPS against 133.7k for 32 rows, one wave).
  const long slots = static_cast<long>(per_cu) * cus;
  const long tiles_x = (nx + block_x - 1) / block_x;
  int best = 16;
  double best_fill = -1.0;
  for (const int rows : {64, 48, 32, 24, 16}) {
    const long tiles = tiles_x * ((ny + rows - 1) / rows);
    const double fill = static_cast<double>(tiles) / (((tiles + slots - 1) / slots) * slots);
    if ((fill >= 0.97 || tiles >= 3 * slots) && (tiles >= min_waves * slots || rows == 16))
      return rows;
    if (fill > best_fill) {
      best_fill = fill;
      best = rows;
    }
  }
  return best;
}

template <class T>
int step2_tile_rows(const StepParams<T>& p, int block_x, int min_waves) {
  const size_t lds = step2_lds_bytes(block_x, sizeof(T));
  const bool one = p.one_rsrc && static_cast<uint64_t>(kQ) * p.N * sizeof(T) < (uint64_t{1} << 32);
  return one ? tile_rows_for(step2_kernel<T, PlanesOne, kAuxNt>, p.nx, p.nrows, block_x, lds,
                             min_waves)
             : tile_rows_for(step2_kernel<T, Planes, kAuxNt>, p.nx, p.nrows, block_x, lds,
                             min_waves);
}
template int step2_tile_rows<float>(const StepParams<float>&, int, int);
template int step2_tile_rows<double>(const StepParams<double>&, int, int);

namespace {
bool fp16_one_resource(const StepParamsH& p) {
  return static_cast<uint64_t>(kQ) * p.N * sizeof(uint16_t) < (uint64_t{1} << 32);
}
}  // namespace

hipError_t launch_step2_fp16(const StepParamsH& p, int block_x, hipStream_t stream) {
  // from an even step (shift 0), periodic, uniform force, whole domain, tile rows >= 1, planes
  // addressable with 32 bits and holding the discard slots
  if (p.shift != 0 || !p.wrap_y || p.force || p.row0 != 0 || p.nrows != p.ny || p.rows_per_tile < 1 ||
      block_x < 64 || block_x > 512 || block_x % 64 ||
      p.tiles_x != (p.nx + block_x - 1) / block_x ||
      static_cast<uint64_t>(p.N) * sizeof(uint16_t) >= (uint64_t{1} << 32) ||
      p.N < static_cast<size_t>(p.nx) * p.ny + kPlaneTail)
    return hipErrorInvalidValue;
  const size_t lds = step2_lds_bytes(block_x, sizeof(_Float16));
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  const unsigned tiles_y = static_cast<unsigned>((p.ny + p.rows_per_tile - 1) / p.rows_per_tile);
  if (p.xcd_bands == 2 && tiles_y % 8) return hipErrorInvalidValue;
  const unsigned grid = static_cast<unsigned>(p.tiles_x) * tiles_y;
  if (fp16_one_resource(p))
    step2_fp16_kernel<PlanesOne><<<grid, block_x, lds, stream>>>(p);
  else
    step2_fp16_kernel<Planes><<<grid, block_x, lds, stream>>>(p);
  return hipGetLastError();
}

int step2_tile_rows_fp16(const StepParamsH& p, int block_x) {
  const size_t lds = step2_lds_bytes(block_x, sizeof(_Float16));
  return fp16_one_resource(p) ? tile_rows_for(step2_fp16_kernel<PlanesOne>, p.nx, p.ny, block_x, lds)
                              : tile_rows_for(step2_fp16_kernel<Planes>, p.nx, p.ny, block_x, lds);
}

}  // namespace lbmhip
