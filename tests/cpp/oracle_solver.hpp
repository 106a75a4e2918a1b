// Test infrastructure: the oracle (oracle/lbm_oracle.cpp, sequential push v0 restatement) behind
// the reference's duck-typed Solver interface, so the drop-in tests can instantiate
// run<OracleV0>() and run<SolverHip>() from one template, as the reference's
// tests/test_v3_vs_v0.cu does with SolverV0 and SolverV3.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <vector>

#include "lbm_hip_solver.hpp"

extern "C" {
void* oracle_create(int kind, int precision, int nx, int ny, double tau, double fx, double fy,
                    const std::uint8_t* solid);
void oracle_destroy(void*);
void oracle_set_inflow_outflow(void*, double);
void oracle_init_equilibrium(void*, double, double, double);
void oracle_run(void*, int);
void oracle_populations(void*, void*);
double oracle_total_mass(void*);
}

namespace lbm {

class OracleV0 {
 public:
  OracleV0(int nx, int ny, real tau, real fx, real fy, Mask solid)
      : nx_(nx), ny_(ny), solid_(std::move(solid)) {
    h_ = oracle_create(0, kHipPrecision, nx, ny, tau, fx, fy, solid_.data());
    if (!h_) throw std::invalid_argument("oracle_create failed");
  }
  ~OracleV0() { oracle_destroy(h_); }
  OracleV0(const OracleV0&) = delete;
  OracleV0& operator=(const OracleV0&) = delete;
  void set_inflow_outflow(real u) { oracle_set_inflow_outflow(h_, u); }
  void init_equilibrium(real r, real u, real v) { oracle_init_equilibrium(h_, r, u, v); }
  void run(int steps) { oracle_run(h_, steps); }
  std::vector<real> populations() const {
    std::vector<real> g(static_cast<std::size_t>(nx_) * ny_ * 9);
    oracle_populations(h_, g.data());
    return g;
  }
  double total_mass() const { return oracle_total_mass(h_); }
  int nx() const { return nx_; }
  int ny() const { return ny_; }
  const Mask& solid() const { return solid_; }

 private:
  int nx_, ny_;
  Mask solid_;
  void* h_;
};

}  // namespace lbm
