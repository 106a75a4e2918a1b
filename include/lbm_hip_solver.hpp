// lbm_hip_solver.hpp — header-only C++17 class with the exact method set of the reference's
// SolverV3 (src/v3_cuda_opt/solver_v3.hpp:27-57), implemented over the C ABI of lbm_hip.h, so
// the reference's run_driver<Solver> (src/common/driver.hpp:255-300) and its test templates
// (tests/test_v3_vs_v0.cu:30-37, tests/poiseuille_check.hpp:81-113, tests/cylinder_case.hpp:
// 23-31) compile unchanged against lbm::SolverHip.
//
// Precision follows the reference's switch (src/common/real.hpp:7-11): real = double when
// LBM_REAL_DOUBLE is defined, float otherwise. Errors follow the reference: invalid arguments
// throw std::invalid_argument, runtime (HIP) failures throw std::runtime_error.
#ifndef LBM_HIP_SOLVER_HPP
#define LBM_HIP_SOLVER_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "lbm_hip.h"

namespace lbm {

#ifdef LBM_REAL_DOUBLE
using real = double;
inline constexpr int kHipPrecision = LBM_DOUBLE;
#else
using real = float;
inline constexpr int kHipPrecision = LBM_FLOAT;
#endif
using Mask = std::vector<std::uint8_t>;

class SolverHip {
 public:
  SolverHip(int nx, int ny, real tau, real force_x, real force_y, Mask solid)
      : nx_(nx), ny_(ny), solid_(std::move(solid)) {
    if (solid_.size() != static_cast<std::size_t>(nx) * static_cast<std::size_t>(ny))
      throw std::invalid_argument("solid mask size does not match the grid");
    check(lbm_create(nx, ny, tau, force_x, force_y, solid_.data(), kHipPrecision, &h_));
  }
  ~SolverHip() { lbm_destroy(h_); }
  SolverHip(const SolverHip&) = delete;
  SolverHip& operator=(const SolverHip&) = delete;

  void init_equilibrium(real rho0, real ux0, real uy0) {
    check(lbm_init_equilibrium(h_, rho0, ux0, uy0));
  }
  void set_inflow_outflow(real u_in) { check(lbm_set_inflow_outflow(h_, u_in)); }
  void step() { check(lbm_step(h_)); }
  void run(int steps) { check(lbm_run(h_, steps)); }
  double timed_run(int steps) {
    double s = 0.0;
    check(lbm_timed_run(h_, steps, &s));
    return s;
  }

  void macroscopic(std::vector<real>& rho, std::vector<real>& ux, std::vector<real>& uy) const {
    const std::size_t n = cells();
    rho.resize(n);
    ux.resize(n);
    uy.resize(n);
    check(lbm_macroscopic(h_, rho.data(), ux.data(), uy.data()));
  }
  double total_mass() const {
    double m = 0.0;
    check(lbm_total_mass(h_, &m));
    return m;
  }
  std::vector<real> populations() const {
    std::vector<real> g(cells() * 9);
    check(lbm_populations(h_, g.data()));
    return g;
  }
  void populations_at(int x, int y, real* out) const { check(lbm_populations_at(h_, x, y, out)); }

  int threads() const { return 0; }
  std::string bench_info() const {
    char buf[1024] = {};
    check(lbm_bench_info(h_, buf, sizeof buf));
    return buf;
  }
  int nx() const { return nx_; }
  int ny() const { return ny_; }
  const Mask& solid() const { return solid_; }

  // Beyond SolverV3: setters, momentum, launch shape, raw handle.
  void set_tau(real tau) { check(lbm_set_tau(h_, tau)); }
  void set_force(real fx, real fy) { check(lbm_set_force(h_, fx, fy)); }
  void set_mask(Mask solid) {
    // the C ABI reads nx * ny bytes (the constructor's check, solver_v3.cu:169-170)
    if (solid.size() != cells()) throw std::invalid_argument("solid mask size does not match the grid");
    check(lbm_set_mask(h_, solid.data()));
    solid_ = std::move(solid);
  }
  void set_block(int block_x, int rows) { check(lbm_set_block(h_, block_x, rows)); }
  std::pair<double, double> momentum() const {
    double px = 0.0, py = 0.0;
    check(lbm_momentum(h_, &px, &py));
    return {px, py};
  }
  lbm_solver* handle() const { return h_; }

 private:
  std::size_t cells() const { return static_cast<std::size_t>(nx_) * static_cast<std::size_t>(ny_); }
  static void check(int rc) {
    if (rc == LBM_OK) return;
    if (rc == LBM_EINVAL) throw std::invalid_argument(lbm_last_error());
    throw std::runtime_error(lbm_last_error());
  }

  int nx_, ny_;
  Mask solid_;
  lbm_solver* h_ = nullptr;
};

}  // namespace lbm

#endif  // LBM_HIP_SOLVER_HPP
