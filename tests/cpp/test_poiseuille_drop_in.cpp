// Plane Poiseuille flow on the C++ drop-in (lbm::SolverHip over the C ABI), shaped like the
// reference's Poiseuille template (tests/poiseuille_check.hpp:81-158, run for v3 by
// tests/test_poiseuille_v3.cu): a periodic channel of nx = 4 columns, walls on rows 0 and
// ny - 1 (halfway bounce-back, height H = ny - 2), driven by the body force that gives a peak
// velocity of 0.05, run from rest in chunks of 1000 steps until the x-averaged u_x profile
// stops changing, then compared with the closed-form parabola. Built in float and in double
// (-DLBM_REAL_DOUBLE) with the reference's tolerances:
//   A  tau = 1/2 + sqrt(3)/4, where BGK with halfway bounce-back is exact: L2 < 1e-8 (double),
//      2e-4 (float: the run stagnates at float resolution);
//   B  tau = 1, H = 16, 32, 64: L2 < 2e-3, and in double an observed order in (1.9, 2.1);
//   C  relative mass drift < 1e-13 (double), 1e-6 (float).
// The solver is a template parameter: any class with SolverV3's method set runs the same
// checks.
#include <cmath>
#include <cstdio>
#include <vector>

#include "lbm_hip_solver.hpp"

namespace {

using namespace lbm;

constexpr bool kDouble = sizeof(real) == sizeof(double);
constexpr double kSteadyTol = kDouble ? 1e-10 : 1e-7;  // relative profile change per chunk
constexpr double kExactTol = kDouble ? 1e-8 : 2e-4;
constexpr double kErrorTol = 2e-3;
constexpr double kMassTol = kDouble ? 1e-13 : 1e-6;
constexpr int kChunk = 1000, kNx = 4;
constexpr double kUmax = 0.05;

int failures = 0;

void expect(bool ok, const char* what) {
  std::printf("  [%s] %s\n", ok ? "PASS" : "FAIL", what);
  if (!ok) ++failures;
}

// Kinematic viscosity and the driving force, in the solver's real like the reference
// (d2q9.hpp:30, geometry.hpp:24-26): u_max = F H^2 / (8 nu).
real viscosity(real tau) { return (tau - real(0.5)) / real(3); }
real driving_force(int H, real tau) {
  return real(8) * viscosity(tau) * real(kUmax) / (real(H) * real(H));
}

Mask channel(int nx, int ny) {
  Mask m(static_cast<std::size_t>(nx) * ny, 0);
  for (int x = 0; x < nx; ++x) m[x] = m[static_cast<std::size_t>(ny - 1) * nx + x] = 1;
  return m;
}

// Mean u_x of each fluid row y = 1 .. ny - 2.
template <class Solver>
std::vector<double> row_means(const Solver& s) {
  std::vector<real> rho, ux, uy;
  s.macroscopic(rho, ux, uy);
  std::vector<double> m(static_cast<std::size_t>(s.ny() - 2), 0.0);
  for (int y = 1; y + 1 < s.ny(); ++y) {
    double sum = 0.0;
    for (int x = 0; x < s.nx(); ++x) sum += ux[static_cast<std::size_t>(y) * s.nx() + x];
    m[y - 1] = sum / s.nx();
  }
  return m;
}

double relative_l2(const std::vector<double>& a, const std::vector<double>& b) {
  double d = 0.0, n = 0.0;
  for (std::size_t k = 0; k < a.size(); ++k) {
    d += (a[k] - b[k]) * (a[k] - b[k]);
    n += b[k] * b[k];
  }
  return std::sqrt(d / n);
}

struct Outcome {
  double l2, drift;
  long steps;
  bool steady;
};

template <class Solver>
Outcome channel_flow(int H, double tau_d) {
  const real tau = real(tau_d);
  const real F = driving_force(H, tau);
  const double nu = static_cast<double>(viscosity(tau));
  Solver s(kNx, H + 2, tau, F, real(0), channel(kNx, H + 2));
  s.init_equilibrium(real(1), real(0), real(0));
  const double m0 = s.total_mass();
  // at most 20 viscous times H^2 / nu
  const long cap = static_cast<long>(20.0 * H * H / nu);
  std::vector<double> last = row_means(s);
  Outcome out{0.0, 0.0, 0, false};
  while (out.steps < cap) {
    s.run(kChunk);
    out.steps += kChunk;
    std::vector<double> now = row_means(s);
    const double change = relative_l2(now, last);
    last = std::move(now);
    if (change < kSteadyTol) {
      out.steady = true;
      break;
    }
  }
  // halfway walls at y = 1/2 and H + 1/2: fluid row k + 1 lies k + 1/2 from the lower one
  std::vector<double> exact(static_cast<std::size_t>(H));
  for (int k = 0; k < H; ++k) {
    const double yw = k + 0.5;
    exact[k] = static_cast<double>(F) / (2.0 * nu) * yw * (H - yw);
  }
  out.l2 = relative_l2(last, exact);
  out.drift = (s.total_mass() - m0) / m0;
  std::printf("  H=%d tau=%.4f L2=%.3e steps=%ld%s mass drift=%.2e\n", H, tau_d, out.l2,
              out.steps, out.steady ? "" : " (not steady)", out.drift);
  return out;
}

template <class Solver>
void poiseuille_checks() {
  std::printf("A. exact tau\n");
  const Outcome a = channel_flow<Solver>(32, 0.5 + std::sqrt(3.0) / 4.0);
  expect(a.l2 < kExactTol, "L2 error at round-off level");
  expect(std::fabs(a.drift) < kMassTol, "mass conserved");
  std::printf("B. tau = 1, H = 16, 32, 64\n");
  const int Hs[3] = {16, 32, 64};
  double e[3];
  for (int k = 0; k < 3; ++k) {
    const Outcome b = channel_flow<Solver>(Hs[k], 1.0);
    e[k] = b.l2;
    expect(b.l2 < kErrorTol, "L2 error below tolerance");
    expect(std::fabs(b.drift) < kMassTol, "mass conserved");
  }
  for (int k = 0; k < 2; ++k) {
    const double order = std::log2(e[k] / e[k + 1]);
    std::printf("  observed order H=%d -> %d: %.3f\n", Hs[k], Hs[k + 1], order);
    if (kDouble) expect(order > 1.9 && order < 2.1, "second-order convergence");
  }
}

}  // namespace

int main() {
  std::printf("Poiseuille on lbm::SolverHip, real = %s\n", kDouble ? "double" : "float");
  poiseuille_checks<SolverHip>();
  std::printf("%s\n", failures == 0 ? "ALL PASSED" : "FAILED");
  return failures == 0 ? 0 : 1;
}
