// The C++ drop-in (lbm::SolverHip over the C ABI) must reproduce v0 bit for bit, shaped like the
// reference's tests/test_v3_vs_v0.cu (walls + block, 53x31, 500 steps, tau 0.8, F = (1e-5,
// 2e-6), init (1, 0.02, 0.01)) and tests/test_cylinder_v3.cu (120x48 cylinder, D = 10, open x,
// Re = 20, 1000 steps); launch shapes are swept like the reference's variants. Built in float
// and in double (-DLBM_REAL_DOUBLE). Also checks that invalid arguments throw
// std::invalid_argument as the reference does.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "lbm_hip_solver.hpp"
#include "oracle_solver.hpp"

namespace {

using namespace lbm;

constexpr int kNx = 53, kNy = 31, kSteps = 500;
constexpr real kTau = real(0.8), kFx = real(1e-5), kFy = real(2e-6);
const char* kRealName = sizeof(real) == sizeof(double) ? "double" : "float";

Mask channel(int nx, int ny) {
  Mask m(static_cast<std::size_t>(nx) * ny, 0);
  for (int x = 0; x < nx; ++x) m[x] = m[static_cast<std::size_t>(ny - 1) * nx + x] = 1;
  return m;
}

Mask walls_and_block() {
  Mask m = channel(kNx, kNy);
  for (int y = 10; y < 17; ++y)
    for (int x = 12; x < 16; ++x) m[static_cast<std::size_t>(y) * kNx + x] = 1;
  return m;
}

Mask cylinder(int nx, int ny, double xc, double yc, double D) {
  Mask m(static_cast<std::size_t>(nx) * ny, 0);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const double dx = x - xc, dy = y - yc;
      if (dx * dx + dy * dy <= 0.25 * D * D) m[static_cast<std::size_t>(y) * nx + x] = 1;
    }
  return m;
}

template <class Solver>
std::vector<real> run_block(int bx = 0, int rows = 0) {
  Solver s(kNx, kNy, kTau, kFx, kFy, walls_and_block());
  if constexpr (std::is_same_v<Solver, SolverHip>)
    if (bx) s.set_block(bx, rows);
  s.init_equilibrium(real(1), real(0.02), real(0.01));
  s.run(kSteps);
  return s.populations();
}

template <class Solver>
std::vector<real> run_cylinder(int bx = 0, int rows = 0) {
  constexpr int nx = 120, ny = 48, D = 10;
  const real u = real(0.05);
  const real tau = static_cast<real>(0.5 + 3.0 * static_cast<double>(u) * D / 20.0);
  Solver s(nx, ny, tau, real(0), real(0), cylinder(nx, ny, 3.0 * D, ny / 2 + 1, D));
  if constexpr (std::is_same_v<Solver, SolverHip>)
    if (bx) s.set_block(bx, rows);
  s.set_inflow_outflow(u);
  s.init_equilibrium(real(1), u, real(0));
  s.run(1000);
  return s.populations();
}

int compare(const char* what, const std::vector<real>& g, const std::vector<real>& ref) {
  std::size_t differing = 0;
  double max_diff = 0.0;
  for (std::size_t k = 0; k < ref.size(); ++k)
    if (std::memcmp(&g[k], &ref[k], sizeof(real)) != 0) {
      ++differing;
      max_diff = std::fmax(max_diff, std::fabs(static_cast<double>(g[k]) - ref[k]));
    }
  std::printf("  [%s] %-28s differing values=%zu / %zu  max |diff|=%.3e\n",
              differing ? "FAIL" : "PASS", what, differing, ref.size(), max_diff);
  return differing ? 1 : 0;
}

}  // namespace

int main() {
  std::printf("drop-in SolverHip vs oracle v0, real = %s\n", kRealName);
  int failures = 0;
  const auto ref = run_block<OracleV0>();
  const int shapes[][2] = {{0, 0}, {64, 1}, {128, 4}, {256, 8}, {1024, 3}};
  for (const auto& sh : shapes) {
    char name[64];
    std::snprintf(name, sizeof name, "walls+block %dx%d", sh[0], sh[1]);
    failures += compare(name, run_block<SolverHip>(sh[0], sh[1]), ref);
  }
  const auto cref = run_cylinder<OracleV0>();
  failures += compare("cylinder open-x", run_cylinder<SolverHip>(), cref);
  failures += compare("cylinder open-x 64x2", run_cylinder<SolverHip>(64, 2), cref);

  bool threw = false;
  try {
    SolverHip bad(kNx, kNy, real(0.5), kFx, kFy, walls_and_block());
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  std::printf("  [%s] tau = 0.5 throws std::invalid_argument\n", threw ? "PASS" : "FAIL");
  failures += threw ? 0 : 1;
  std::printf("%s\n", failures == 0 ? "ALL PASSED" : "FAILED");
  return failures == 0 ? 0 : 1;
}
