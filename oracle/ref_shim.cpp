// ORACLE — test infrastructure only. A C-ABI shim over the reference's own CPU solvers (v0, v1)
// and its field writers (io.cpp: write_npy, vorticity),
// compiled from the sources where they lie under the reference checkout by oracle/Makefile
// (targets ref-float / ref-double) into oracle/_ref/. No reference source is copied here: this
// file only includes the reference headers and forwards to SolverV0 / SolverV1.
//
// Used to pin the oracle restatement (tests/golden/make_golden.py writes the reference's own
// outputs as fixtures) and, when present, as the "reference" CPU baseline of bench.py.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

#include "geometry.hpp"
#include "io.hpp"
#include "real.hpp"
#include "solver_v0.hpp"
#include "solver_v1.hpp"

using lbm::real;

namespace {

template <class Solver>
void run_into(int nx, int ny, double tau, double fx, double fy, const std::uint8_t* solid,
              int open_x, double u_in, double rho0, double ux0, double uy0, int steps,
              real* out) {
  lbm::Mask m(solid, solid + static_cast<std::size_t>(nx) * ny);
  Solver s(nx, ny, static_cast<real>(tau), static_cast<real>(fx), static_cast<real>(fy), m);
  if (open_x) s.set_inflow_outflow(static_cast<real>(u_in));
  s.init_equilibrium(static_cast<real>(rho0), static_cast<real>(ux0), static_cast<real>(uy0));
  s.run(steps);
  const auto& g = s.populations();
  std::memcpy(out, g.data(), g.size() * sizeof(real));
}

}  // namespace

extern "C" {

int ref_real_size() { return static_cast<int>(sizeof(real)); }

// Post-streaming populations (v0 AoS layout) after `steps` steps of SolverV0.
void ref_v0_run(int nx, int ny, double tau, double fx, double fy, const std::uint8_t* solid,
                int open_x, double u_in, double rho0, double ux0, double uy0, int steps,
                void* out) {
  run_into<lbm::SolverV0>(nx, ny, tau, fx, fy, solid, open_x, u_in, rho0, ux0, uy0, steps,
                          static_cast<real*>(out));
}

void ref_v1_run(int nx, int ny, double tau, double fx, double fy, const std::uint8_t* solid,
                int open_x, double u_in, double rho0, double ux0, double uy0, int steps,
                void* out) {
  run_into<lbm::SolverV1>(nx, ny, tau, fx, fy, solid, open_x, u_in, rho0, ux0, uy0, steps,
                          static_cast<real*>(out));
}

double ref_v0_total_mass(int nx, int ny, double tau, double fx, double fy,
                         const std::uint8_t* solid, double rho0, double ux0, double uy0,
                         int steps) {
  lbm::Mask m(solid, solid + static_cast<std::size_t>(nx) * ny);
  lbm::SolverV0 s(nx, ny, static_cast<real>(tau), static_cast<real>(fx), static_cast<real>(fy), m);
  s.init_equilibrium(static_cast<real>(rho0), static_cast<real>(ux0), static_cast<real>(uy0));
  s.run(steps);
  return s.total_mass();
}

void ref_make_cylinder(int nx, int ny, double xc, double yc, double D, std::uint8_t* out) {
  const lbm::Mask m = lbm::make_cylinder(nx, ny, xc, yc, D);
  std::memcpy(out, m.data(), m.size());
}

// CPU baseline: a walled channel (make_channel) run with SolverV1 after `warmup` untimed
// steps; returns the seconds of `steps` timed steps (the host clock, as driver.hpp:90-99).
double ref_v1_channel_timed(int nx, int ny, double tau, double fx, int warmup, int steps,
                            int* threads) {
  lbm::SolverV1 s(nx, ny, static_cast<real>(tau), static_cast<real>(fx), real(0),
                  lbm::make_channel(nx, ny));
  s.init_equilibrium(real(1), real(0), real(0));
  s.run(warmup);
  const auto t0 = std::chrono::steady_clock::now();
  s.run(steps);
  const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (threads) *threads = s.threads();
  return t;
}

// The reference's timing protocol (src/common/driver.hpp:115-132) on SolverV1: one solver on
// the given mask, rest initial state, `warmup` untimed steps, then `reps` timed runs of `steps`
// steps (host clock around run(), as driver.hpp's timed() does for a solver without
// timed_run); seconds[r] receives run r's time.
void ref_v1_protocol(int nx, int ny, double tau, double fx, double fy, const std::uint8_t* solid,
                     int warmup, int steps, int reps, double* seconds, int* threads) {
  lbm::Mask m(solid, solid + static_cast<std::size_t>(nx) * ny);
  lbm::SolverV1 s(nx, ny, static_cast<real>(tau), static_cast<real>(fx), static_cast<real>(fy), m);
  s.init_equilibrium(real(1), real(0), real(0));
  s.run(warmup);
  for (int r = 0; r < reps; ++r) {
    const auto t0 = std::chrono::steady_clock::now();
    s.run(steps);
    seconds[r] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  if (threads) *threads = s.threads();
}

// io.cpp's vorticity (src/common/io.cpp:54-74) of caller-owned ux, uy (nx * ny reals).
void ref_vorticity(const void* ux, const void* uy, const std::uint8_t* solid, int nx, int ny,
                   int periodic_x, void* out) {
  const std::size_t n = static_cast<std::size_t>(nx) * ny;
  const real* u = static_cast<const real*>(ux);
  const real* v = static_cast<const real*>(uy);
  const std::vector<real> a(u, u + n), b(v, v + n);
  const lbm::Mask m(solid, solid + n);
  const std::vector<real> w = lbm::vorticity(a, b, m, nx, ny, periodic_x != 0);
  std::memcpy(out, w.data(), n * sizeof(real));
}

// io.cpp's write_npy (src/common/io.cpp:11-52): kind 0 the solver real, 1 uint8.
void ref_write_npy(const char* path, const void* a, int ny, int nx, int kind) {
  const std::size_t n = static_cast<std::size_t>(nx) * ny;
  if (kind == 1) {
    const std::uint8_t* p = static_cast<const std::uint8_t*>(a);
    lbm::write_npy(path, std::vector<std::uint8_t>(p, p + n), ny, nx);
  } else {
    const real* p = static_cast<const real*>(a);
    lbm::write_npy(path, std::vector<real>(p, p + n), ny, nx);
  }
}

}  // extern "C"
