// ORACLE — test infrastructure only. Nothing in the product library links, loads or calls this
// file; only tests/, the smoke check and the CPU-baseline leg of bench.py use it, as a checker.
//
// CPU restatement of the reference solver's D2Q9 time step (BGK collision, Guo body force,
// halfway bounce-back, populations stored as g = f - w), in two forms:
//
//   kind 0, "v0": sequential push on an array-of-structures state g[(y*nx + x)*9 + i], two
//       buffers and a swap. Follows src/v0_cpu/solver_v0.cpp:63-135 operation for operation
//       (moments :26-39, collision :76-83, push with periodic wrap and bounce-back :85-99,
//       open-x rules src/common/boundary.hpp:37-48), so a float or double run of it is the
//       bitwise reference for every other path.
//   kind 1, "pull": row-parallel OpenMP pull on a structure-of-arrays state R[i*N + cell]
//       holding post-collision populations; the CPU baseline timed beside the GPU. Structure of
//       src/v1_omp/solver_v1.cpp:136-175 (rows split statically, periodic rows computed once
//       per row, wrap columns apart, both loads then a select, solid cells stored as zero).
//       Its post-streaming view (pull_to_aos) must equal kind 0 bit for bit.
//
// The numerics (equilibrium, Guo source, moments) follow src/common/d2q9.hpp:40-50 with the
// same parse order; the tables follow src/common/d2q9.hpp:22-27. Built with -ffp-contract=off
// and without fast-math, like the reference (CMakeLists.txt:18-24).
//
// Extension beyond the reference: an optional per-cell body force field (planar fx then fy),
// used where the reference passes its uniform (fx, fy); with a uniform field the arithmetic is
// identical to the uniform case.
#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

namespace oracle {

constexpr int Q = 9;
// Directions: 0 rest, 1 E, 2 N, 3 W, 4 S, 5 NE, 6 NW, 7 SW, 8 SE (src/common/d2q9.hpp:14-27).
constexpr int CX[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int OPP[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

template <class T>
constexpr T weight(int i) {
  return i == 0 ? T(4) / T(9) : (i <= 4 ? T(1) / T(9) : T(1) / T(36));
}

// Shifted equilibrium f_eq - w (src/common/d2q9.hpp:40-42), same parse order.
template <class T>
inline T equilibrium(T wi, T drho, T rho, T cu, T usq) {
  return wi * (drho + rho * (T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq));
}

// Guo source without its (1 - omega/2) prefactor (src/common/d2q9.hpp:46-50).
template <class T>
inline T guo(T wi, T cxi, T cyi, T ux, T uy, T fx, T fy) {
  const T cu = cxi * ux + cyi * uy;
  const T cf = cxi * fx + cyi * fy;
  return wi * (T(3) * ((cxi - ux) * fx + (cyi - uy) * fy) + T(9) * cu * cf);
}

// Moments of nine populations: m0 = sum g, m = sum c g, summed i = 0..8 from +0
// (src/v0_cpu/solver_v0.cpp:26-39).
template <class T>
inline void moments(const T* gi, T fx, T fy, T& drho, T& ux, T& uy) {
  T m0 = 0, mx = 0, my = 0;
  for (int i = 0; i < Q; ++i) {
    m0 += gi[i];
    mx += CX[i] * gi[i];
    my += CY[i] * gi[i];
  }
  drho = m0;
  const T rho = T(1) + drho;
  ux = (mx + T(0.5) * fx) / rho;
  uy = (my + T(0.5) * fy) / rho;
}

// BGK + Guo collision of nine gathered populations, in place (src/v0_cpu/solver_v0.cpp:72-83).
template <class T>
inline void collide(T* gi, T fx, T fy, T omega, T fp) {
  T drho, ux, uy;
  moments(gi, fx, fy, drho, ux, uy);
  const T rho = T(1) + drho;
  const T usq = ux * ux + uy * uy;
  T out[Q];
  for (int i = 0; i < Q; ++i) {
    const T cu = CX[i] * ux + CY[i] * uy;
    const T geq = equilibrium(weight<T>(i), drho, rho, cu, usq);
    out[i] = gi[i] - omega * (gi[i] - geq) +
             fp * guo(weight<T>(i), T(CX[i]), T(CY[i]), ux, uy, fx, fy);
  }
  for (int i = 0; i < Q; ++i) gi[i] = out[i];
}

// Uniform equilibrium of (rho0, ux0, uy0) (src/common/boundary.hpp:26-34).
template <class T>
inline void equilibrium_state(T rho0, T ux0, T uy0, T* eq) {
  const T usq = ux0 * ux0 + uy0 * uy0;
  const T drho0 = rho0 - T(1);
  for (int i = 0; i < Q; ++i) eq[i] = equilibrium(weight<T>(i), drho0, rho0, CX[i] * ux0 + CY[i] * uy0, usq);
}

// Post-streaming populations of fluid cell (x, y) from a pull (post-collision, SoA) state
// (src/common/boundary.hpp:53-70): inlet slots take eq_in, outlet slots copy x = nx - 2,
// bounce-back from a solid source returns the cell's own opposite population.
template <class T>
inline void gather_pull(const T* R, std::size_t N, const std::uint8_t* solid, int nx, int ny, int x,
                        int y, bool open_x, const T* eq_in, T* out) {
  for (int i = 0; i < Q; ++i) {
    if (open_x && x == 0 && CX[i] == 1) {
      out[i] = eq_in[i];
      continue;
    }
    const int xc = (open_x && x == nx - 1 && CX[i] == -1) ? nx - 2 : x;
    const std::size_t cc = static_cast<std::size_t>(y) * nx + xc;
    const int xs = (xc - CX[i] + nx) % nx;
    const int ys = (y - CY[i] + ny) % ny;
    const std::size_t s = static_cast<std::size_t>(ys) * nx + xs;
    out[i] = solid[s] ? R[OPP[i] * N + cc] : R[i * N + s];
  }
}

// ---------------------------------------------------------------------------------------------
// kind 0: sequential push, AoS (src/v0_cpu/solver_v0.cpp).
template <class T>
struct PushAoS {
  int nx, ny;
  T tau, fx, fy;
  std::vector<std::uint8_t> solid;
  std::vector<T> force;  // optional per-cell field, 2N planar
  bool open_x = false;
  T eq_in[Q] = {};
  std::vector<T> g, gn;

  std::size_t cells() const { return static_cast<std::size_t>(nx) * ny; }
  T fxc(std::size_t c) const { return force.empty() ? fx : force[c]; }
  T fyc(std::size_t c) const { return force.empty() ? fy : force[cells() + c]; }

  void init_equilibrium(T rho0, T ux0, T uy0) {
    T eq[Q];
    equilibrium_state(rho0, ux0, uy0, eq);
    for (std::size_t c = 0; c < cells(); ++c)
      for (int i = 0; i < Q; ++i) g[c * Q + i] = solid[c] ? T(0) : eq[i];
    gn = g;
  }

  void step() {
    const T omega = T(1) / tau;
    const T fp = T(1) - T(0.5) * omega;
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const std::size_t c = static_cast<std::size_t>(y) * nx + x;
        if (solid[c]) continue;
        T gi[Q];
        for (int i = 0; i < Q; ++i) gi[i] = g[c * Q + i];
        collide(gi, fxc(c), fyc(c), omega, fp);
        for (int i = 0; i < Q; ++i) {
          int xn = x + CX[i];
          if (open_x) {
            if (xn < 0 || xn >= nx) continue;  // leaves the open domain
          } else {
            xn = (xn + nx) % nx;
          }
          const int yn = (y + CY[i] + ny) % ny;
          const std::size_t cn = static_cast<std::size_t>(yn) * nx + xn;
          if (solid[cn])
            gn[c * Q + OPP[i]] = gi[i];
          else
            gn[cn * Q + i] = gi[i];
        }
      }
    if (open_x) {
      for (int y = 0; y < ny; ++y) {
        const std::size_t row = static_cast<std::size_t>(y) * nx;
        if (!solid[row])
          for (int i : {1, 5, 8}) gn[row * Q + i] = eq_in[i];
        const std::size_t out = row + nx - 1, prev = row + nx - 2;
        if (!solid[out])
          for (int i : {3, 6, 7}) gn[out * Q + i] = gn[prev * Q + i];
      }
    }
    g.swap(gn);
  }

  void populations(T* out) const { std::memcpy(out, g.data(), g.size() * sizeof(T)); }
  void set_populations(const T* in) {
    std::memcpy(g.data(), in, g.size() * sizeof(T));
    gn = g;
  }
};

// ---------------------------------------------------------------------------------------------
// kind 1: OpenMP pull, SoA (structure of src/v1_omp/solver_v1.cpp:18-175).
template <class T>
struct PullSoA {
  int nx, ny;
  T tau, fx, fy;
  std::vector<std::uint8_t> solid;
  std::vector<T> force;
  bool open_x = false;
  T eq_in[Q] = {};
  T* g = nullptr;
  T* gn = nullptr;

  std::size_t cells() const { return static_cast<std::size_t>(nx) * ny; }

  void allocate() {
    const std::size_t n = cells() * Q;
    g = new T[n];
    gn = new T[n];
    const std::size_t N = cells();
    // First touch with the row distribution of step().
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ny; ++y)
      for (int i = 0; i < Q; ++i) {
        T* a = g + i * N + static_cast<std::size_t>(y) * nx;
        T* b = gn + i * N + static_cast<std::size_t>(y) * nx;
        for (int x = 0; x < nx; ++x) a[x] = b[x] = T(0);
      }
  }
  ~PullSoA() {
    delete[] g;
    delete[] gn;
  }

  // Preimage by streaming of the uniform equilibrium (src/v1_omp/solver_v1.cpp:170-197).
  void init_equilibrium(T rho0, T ux0, T uy0) {
    T eq[Q];
    equilibrium_state(rho0, ux0, uy0, eq);
    const std::size_t N = cells();
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const std::size_t c = static_cast<std::size_t>(y) * nx + x;
        for (int i = 0; i < Q; ++i) {
          const bool leaves = open_x && (x + CX[i] < 0 || x + CX[i] >= nx);
          const int xn = (x + CX[i] + nx) % nx;
          const int yn = (y + CY[i] + ny) % ny;
          const std::size_t cn = static_cast<std::size_t>(yn) * nx + xn;
          const T v = solid[c] ? T(0) : (!leaves && solid[cn] ? eq[OPP[i]] : eq[i]);
          g[i * N + c] = v;
          gn[i * N + c] = v;
        }
      }
  }

  static inline void collide_store(T* gi, T* __restrict__ out, const std::uint8_t* __restrict__ sol,
                                   std::size_t N, std::size_t c, T omega, T fp, T fxv, T fyv) {
    collide(gi, fxv, fyv, omega, fp);
    for (int i = 0; i < Q; ++i) out[i * N + c] = sol[c] ? T(0) : gi[i];
  }

  void step() {
    const T omega = T(1) / tau;
    const T fp = T(1) - T(0.5) * omega;
    const int nxl = nx, nyl = ny;
    const std::size_t N = cells();
    const T* __restrict__ src = g;
    T* __restrict__ dst = gn;
    const std::uint8_t* __restrict__ sol = solid.data();
    const T* ff = force.empty() ? nullptr : force.data();
    const T fxu = fx, fyu = fy;
    const bool open = open_x;

#pragma omp parallel for schedule(static)
    for (int y = 0; y < nyl; ++y) {
      const std::size_t row[3] = {static_cast<std::size_t>(y == 0 ? nyl - 1 : y - 1) * nxl,
                                  static_cast<std::size_t>(y) * nxl,
                                  static_cast<std::size_t>(y == nyl - 1 ? 0 : y + 1) * nxl};
      auto cell = [&](int xm, int x, int xp) __attribute__((always_inline)) {
        const std::size_t c = row[1] + x;
        const int col[3] = {xm, x, xp};
        T gi[Q];
        for (int i = 0; i < Q; ++i) {
          const std::size_t s = row[1 - CY[i]] + col[1 - CX[i]];
          const T streamed = src[i * N + s];
          const T bounced = src[OPP[i] * N + c];
          gi[i] = sol[s] ? bounced : streamed;
        }
        collide_store(gi, dst, sol, N, c, omega, fp, ff ? ff[c] : fxu, ff ? ff[N + c] : fyu);
      };
      auto edge = [&](int x) {
        const std::size_t c = row[1] + x;
        T gi[Q];
        gather_pull(src, N, sol, nxl, nyl, x, y, true, eq_in, gi);
        collide_store(gi, dst, sol, N, c, omega, fp, ff ? ff[c] : fxu, ff ? ff[N + c] : fyu);
      };
      if (open)
        edge(0);
      else
        cell(nxl - 1, 0, nxl > 1 ? 1 : 0);
#pragma GCC ivdep
      for (int x = 1; x < nxl - 1; ++x) cell(x - 1, x, x + 1);
      if (nxl > 1) {
        if (open)
          edge(nxl - 1);
        else
          cell(nxl - 2, nxl - 1, 0);
      }
    }
    T* t = g;
    g = gn;
    gn = t;
  }

  void populations(T* out) const {
    const std::size_t N = cells();
    std::memset(out, 0, N * Q * sizeof(T));
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const std::size_t c = static_cast<std::size_t>(y) * nx + x;
        if (!solid[c]) gather_pull(g, N, solid.data(), nx, ny, x, y, open_x, eq_in, out + c * Q);
      }
  }
};

// Preimage of a post-streaming AoS state S: slot (c, i) of R is streamed to (c + e_i, i), or to
// (c, opp i) when c + e_i is solid; a slot leaving an open domain is never read (set to 0).
template <class T>
void aos_to_pull(int nx, int ny, const std::uint8_t* solid, bool open_x, const T* S, T* R) {
  const std::size_t N = static_cast<std::size_t>(nx) * ny;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const std::size_t c = static_cast<std::size_t>(y) * nx + x;
      for (int i = 0; i < Q; ++i) {
        if (solid[c]) {
          R[i * N + c] = T(0);
          continue;
        }
        const int xr = x + CX[i];
        if (open_x && (xr < 0 || xr >= nx)) {
          R[i * N + c] = T(0);
          continue;
        }
        const int xn = (xr + nx) % nx, yn = (y + CY[i] + ny) % ny;
        const std::size_t cn = static_cast<std::size_t>(yn) * nx + xn;
        R[i * N + c] = solid[cn] ? S[c * Q + OPP[i]] : S[cn * Q + i];
      }
    }
}

struct Handle {
  int kind, precision;
  void* impl;
};

template <class T>
struct Ops {
  template <class F>
  static void with(Handle* h, F&& f) {
    if (h->kind == 0)
      f(*static_cast<PushAoS<T>*>(h->impl));
    else
      f(*static_cast<PullSoA<T>*>(h->impl));
  }
};

template <class F>
void dispatch(Handle* h, F&& f) {
  if (h->precision == 0)
    Ops<float>::with(h, f);
  else
    Ops<double>::with(h, f);
}

template <class T, class S>
void fill_common(S& s, int nx, int ny, double tau, double fx, double fy, const std::uint8_t* solid) {
  s.nx = nx;
  s.ny = ny;
  s.tau = static_cast<T>(tau);
  s.fx = static_cast<T>(fx);
  s.fy = static_cast<T>(fy);
  s.solid.assign(solid, solid + static_cast<std::size_t>(nx) * ny);
}

}  // namespace oracle

using namespace oracle;

extern "C" {

void* oracle_create(int kind, int precision, int nx, int ny, double tau, double fx, double fy,
                    const std::uint8_t* solid) {
  if (nx < 1 || ny < 1 || !solid || (kind != 0 && kind != 1) || (precision != 0 && precision != 1))
    return nullptr;
  auto* h = new Handle{kind, precision, nullptr};
  const std::size_t n = static_cast<std::size_t>(nx) * ny * Q;
  auto make = [&](auto tag) {
    using T = decltype(tag);
    if (kind == 0) {
      auto* s = new PushAoS<T>();
      fill_common<T>(*s, nx, ny, tau, fx, fy, solid);
      s->g.assign(n, T(0));
      s->gn.assign(n, T(0));
      h->impl = s;
    } else {
      auto* s = new PullSoA<T>();
      fill_common<T>(*s, nx, ny, tau, fx, fy, solid);
      s->allocate();
      h->impl = s;
    }
  };
  if (precision == 0)
    make(float{});
  else
    make(double{});
  return h;
}

void oracle_destroy(void* hv) {
  auto* h = static_cast<Handle*>(hv);
  if (!h) return;
  dispatch(h, [](auto& s) { delete &s; });
  delete h;
}

void oracle_set_inflow_outflow(void* hv, double u_in) {
  dispatch(static_cast<Handle*>(hv), [&](auto& s) {
    using T = std::remove_reference_t<decltype(s.tau)>;
    s.open_x = true;
    equilibrium_state(T(1), static_cast<T>(u_in), T(0), s.eq_in);
  });
}

void oracle_set_force_field(void* hv, const void* fxy) {
  dispatch(static_cast<Handle*>(hv), [&](auto& s) {
    using T = std::remove_reference_t<decltype(s.tau)>;
    const T* p = static_cast<const T*>(fxy);
    s.force.assign(p, p + 2 * s.cells());
  });
}

void oracle_init_equilibrium(void* hv, double rho0, double ux0, double uy0) {
  dispatch(static_cast<Handle*>(hv), [&](auto& s) {
    using T = std::remove_reference_t<decltype(s.tau)>;
    s.init_equilibrium(static_cast<T>(rho0), static_cast<T>(ux0), static_cast<T>(uy0));
  });
}

// Post-streaming AoS state (v0 layout); the pull kind stores its preimage.
void oracle_set_populations(void* hv, const void* S) {
  auto* h = static_cast<Handle*>(hv);
  dispatch(h, [&](auto& s) {
    using T = std::remove_reference_t<decltype(s.tau)>;
    if constexpr (std::is_same_v<std::remove_reference_t<decltype(s)>, PushAoS<T>>) {
      s.set_populations(static_cast<const T*>(S));
    } else {
      aos_to_pull(s.nx, s.ny, s.solid.data(), s.open_x, static_cast<const T*>(S), s.g);
      std::memcpy(s.gn, s.g, s.cells() * Q * sizeof(T));
    }
  });
}

void oracle_run(void* hv, int steps) {
  dispatch(static_cast<Handle*>(hv), [&](auto& s) {
    for (int k = 0; k < steps; ++k) s.step();
  });
}

double oracle_timed_run(void* hv, int steps) {
  const auto t0 = std::chrono::steady_clock::now();
  oracle_run(hv, steps);
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void oracle_populations(void* hv, void* out) {
  dispatch(static_cast<Handle*>(hv), [&](auto& s) {
    using T = std::remove_reference_t<decltype(s.tau)>;
    s.populations(static_cast<T*>(out));
  });
}

// Raw pull state R (kind 1 only): 9N values, SoA.
int oracle_raw_state(void* hv, void* out) {
  auto* h = static_cast<Handle*>(hv);
  if (h->kind != 1) return -1;
  dispatch(h, [&](auto& s) {
    using T = std::remove_reference_t<decltype(s.tau)>;
    if constexpr (!std::is_same_v<std::remove_reference_t<decltype(s)>, PushAoS<T>>)
      std::memcpy(out, s.g, s.cells() * Q * sizeof(T));
  });
  return 0;
}

// rho, ux, uy from the post-streaming state (src/v0_cpu/solver_v0.cpp:110-122); solid cells 0.
void oracle_macroscopic(void* hv, void* rho, void* ux, void* uy) {
  dispatch(static_cast<Handle*>(hv), [&](auto& s) {
    using T = std::remove_reference_t<decltype(s.tau)>;
    const std::size_t n = s.cells();
    std::vector<T> S(n * Q);
    s.populations(S.data());
    T* r = static_cast<T*>(rho);
    T* u = static_cast<T*>(ux);
    T* v = static_cast<T*>(uy);
    for (std::size_t c = 0; c < n; ++c) {
      r[c] = u[c] = v[c] = T(0);
      if (s.solid[c]) continue;
      const T fxv = s.force.empty() ? s.fx : s.force[c];
      const T fyv = s.force.empty() ? s.fy : s.force[n + c];
      T drho;
      moments(&S[c * Q], fxv, fyv, drho, u[c], v[c]);
      r[c] = T(1) + drho;
    }
  });
}

// Fluid-cell count + sum of g over fluid cells, in double, cell then direction
// (src/v0_cpu/solver_v0.cpp:125-135).
double oracle_total_mass(void* hv) {
  double result = 0.0;
  dispatch(static_cast<Handle*>(hv), [&](auto& s) {
    using T = std::remove_reference_t<decltype(s.tau)>;
    const std::size_t n = s.cells();
    std::vector<T> S(n * Q);
    s.populations(S.data());
    double drho = 0.0;
    std::size_t fluid = 0;
    for (std::size_t c = 0; c < n; ++c) {
      if (s.solid[c]) continue;
      ++fluid;
      for (int i = 0; i < Q; ++i) drho += S[c * Q + i];
    }
    result = static_cast<double>(fluid) + drho;
  });
  return result;
}

// Population momentum sum_c sum_i c_i g_i over fluid cells, in double (no half-force term).
void oracle_momentum(void* hv, double* px, double* py) {
  dispatch(static_cast<Handle*>(hv), [&](auto& s) {
    using T = std::remove_reference_t<decltype(s.tau)>;
    const std::size_t n = s.cells();
    std::vector<T> S(n * Q);
    s.populations(S.data());
    double mx = 0.0, my = 0.0;
    for (std::size_t c = 0; c < n; ++c) {
      if (s.solid[c]) continue;
      for (int i = 0; i < Q; ++i) {
        mx += CX[i] * static_cast<double>(S[c * Q + i]);
        my += CY[i] * static_cast<double>(S[c * Q + i]);
      }
    }
    *px = mx;
    *py = my;
  });
}

int oracle_threads(void* hv) {
  return static_cast<Handle*>(hv)->kind == 0 ? 1 : omp_get_max_threads();
}

void oracle_set_threads(int n) { omp_set_num_threads(n); }

// ---- primitives (for the closed-form and invariant checks) ----
double oracle_equilibrium_f64(double wi, double drho, double rho, double cu, double usq) {
  return equilibrium(wi, drho, rho, cu, usq);
}
float oracle_equilibrium_f32(float wi, float drho, float rho, float cu, float usq) {
  return equilibrium(wi, drho, rho, cu, usq);
}
double oracle_guo_f64(double wi, double cxi, double cyi, double ux, double uy, double fx, double fy) {
  return guo(wi, cxi, cyi, ux, uy, fx, fy);
}
void oracle_collide_f64(double* g9, double fx, double fy, double omega, double fp) {
  collide(g9, fx, fy, omega, fp);
}
void oracle_collide_f32(float* g9, float fx, float fy, float omega, float fp) {
  collide(g9, fx, fy, omega, fp);
}
double oracle_weight_f64(int i) { return weight<double>(i); }
float oracle_weight_f32(int i) { return weight<float>(i); }

void oracle_pull_to_aos(int precision, int nx, int ny, const std::uint8_t* solid, int open_x,
                        double u_in, const void* R, void* S) {
  const std::size_t N = static_cast<std::size_t>(nx) * ny;
  auto run = [&](auto tag) {
    using T = decltype(tag);
    T eq_in[Q] = {};
    if (open_x) equilibrium_state(T(1), static_cast<T>(u_in), T(0), eq_in);
    const T* r = static_cast<const T*>(R);
    T* s = static_cast<T*>(S);
    std::memset(s, 0, N * Q * sizeof(T));
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const std::size_t c = static_cast<std::size_t>(y) * nx + x;
        if (!solid[c]) gather_pull(r, N, solid, nx, ny, x, y, open_x != 0, eq_in, s + c * Q);
      }
  };
  if (precision == 0)
    run(float{});
  else
    run(double{});
}

void oracle_aos_to_pull(int precision, int nx, int ny, const std::uint8_t* solid, int open_x,
                        const void* S, void* R) {
  if (precision == 0)
    aos_to_pull(nx, ny, solid, open_x != 0, static_cast<const float*>(S), static_cast<float*>(R));
  else
    aos_to_pull(nx, ny, solid, open_x != 0, static_cast<const double*>(S), static_cast<double*>(R));
}

// Channel walls at y = 0 and y = ny - 1 (src/common/geometry.cpp:5-12).
void oracle_make_channel(int nx, int ny, std::uint8_t* solid) {
  std::memset(solid, 0, static_cast<std::size_t>(nx) * ny);
  for (int x = 0; x < nx; ++x) {
    solid[x] = 1;
    solid[static_cast<std::size_t>(ny - 1) * nx + x] = 1;
  }
}

// Staircase disc (x - xc)^2 + (y - yc)^2 <= D^2 / 4 in double (src/common/geometry.cpp:14-23).
void oracle_make_cylinder(int nx, int ny, double xc, double yc, double D, std::uint8_t* solid) {
  const double r2 = 0.25 * D * D;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const double dx = x - xc, dy = y - yc;
      solid[static_cast<std::size_t>(y) * nx + x] = (dx * dx + dy * dy <= r2) ? 1 : 0;
    }
}

}  // extern "C"
