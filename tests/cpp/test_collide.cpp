// The library's collide() (csrc/d2q9.hpp) against the reference's collision written out in
// full (src/v0_cpu/solver_v0.cpp:26-39,72-83 with src/common/d2q9.hpp:40-50: every product by
// a lattice constant, zeros included, in the reference's order), bit for bit, in float and
// double. Inputs: f - w populations at physical scale, signed zeros, subnormals, mixed
// magnitudes; forces of both signs, +0 and -0; tau over (0.505, 2.005), and in a second run
// also the largest tau below 2 and populations with a -0 next to a subnormal (where an
// equilibrium can underflow to -0). Cases whose reference result is not finite are counted and skipped (no bitwise
// counterpart). Host code only: no GPU is used.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>

#include "d2q9.hpp"

namespace {

using lbmhip::kCx;
using lbmhip::kCy;
using lbmhip::kQ;

template <class T>
void reference_collide(T (&g)[kQ], T fx, T fy, T omega, T fp) {
  T m0 = T(0), mx = T(0), my = T(0);
  for (int i = 0; i < kQ; ++i) {
    m0 += g[i];
    mx += T(kCx[i]) * g[i];
    my += T(kCy[i]) * g[i];
  }
  const T drho = m0;
  const T rho = T(1) + drho;
  const T ux = (mx + T(0.5) * fx) / rho;
  const T uy = (my + T(0.5) * fy) / rho;
  const T usq = ux * ux + uy * uy;
  for (int i = 0; i < kQ; ++i) {
    const T cxi = T(kCx[i]), cyi = T(kCy[i]), wi = lbmhip::weight<T>(i);
    const T cu = cxi * ux + cyi * uy;
    const T geq = wi * (drho + rho * (T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq));
    const T cf = cxi * fx + cyi * fy;
    const T src = wi * (T(3) * ((cxi - ux) * fx + (cyi - uy) * fy) + T(9) * cu * cf);
    g[i] = g[i] - omega * (g[i] - geq) + fp * src;
  }
}

struct Rng {
  uint64_t s;
  uint64_t next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {
    const double u = uniform() + 1e-300, v = uniform();
    return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
  }
};

template <class T>
T draw(Rng& r, int kind) {
  const T sign = (r.next() & 1) ? T(-1) : T(1);
  switch (kind) {
    case 0: return static_cast<T>(r.normal() * 1e-2);                      // physical f - w
    case 1: return sign * T(0);                                            // signed zero
    case 2: return sign * std::numeric_limits<T>::denorm_min() *           // subnormal
                   static_cast<T>(1 + (r.next() % 1000));
    case 3: return static_cast<T>(r.normal() * std::ldexp(1.0, -100));     // tiny normal
    case 4: return static_cast<T>(r.normal() * 1e-6);                      // near rest
    default: return static_cast<T>(r.normal() * 10.0);                     // large
  }
}

template <class T>
bool same_bits(T a, T b) {
  return std::memcmp(&a, &b, sizeof(T)) == 0;
}

// Largest tau below 2 whose omega = 1/tau is still above 1/2 in T.
template <class T>
T tau_below_two() {
  T t = std::nextafter(T(2), T(0));
  while (!(T(1) / t > T(0.5))) t = std::nextafter(t, T(0));
  return t;
}

template <class T, bool kEdge>
int run(const char* name, long cases) {
  Rng r{kEdge ? 0x2545F4914F6CDD1Dull : 0x9E3779B97F4A7C15ull};
  long compared = 0, skipped = 0, bad = 0;
  for (long n = 0; n < cases; ++n) {
    const int mode = static_cast<int>(n % 7);  // 6: every population of its own kind
    T g[kQ], h[kQ];
    for (int i = 0; i < kQ; ++i) {
      const int kind = mode == 6 ? static_cast<int>(r.next() % 6) : mode;
      g[i] = draw<T>(r, kind);
      h[i] = g[i];
    }
    const int fk = static_cast<int>(r.next() % 5);
    auto force = [&](int k) -> T {
      return k == 0 ? T(0) : k == 1 ? -T(0) : k == 2 ? static_cast<T>(1e-6)
                                                    : k == 3 ? static_cast<T>(-3e-6)
                                                             : static_cast<T>(r.normal() * 1e-4);
    };
    const T fx = force(fk), fy = force(static_cast<int>(r.next() % 5));
    T tau = static_cast<T>(0.505 + 1.5 * r.uniform());
    if (kEdge) {
      // 1 in 8: tau just below 2; 1 in 16 more: one population -0 with a subnormal neighbour
      const uint64_t pick = r.next() % 16;
      if (pick < 2) tau = tau_below_two<T>();
      else if (tau >= T(2)) tau = static_cast<T>(1.0 + r.uniform());
      if (pick == 2) {
        const int i = static_cast<int>(r.next() % kQ);
        g[i] = h[i] = -T(0);
        g[(i + 1) % kQ] = h[(i + 1) % kQ] = std::numeric_limits<T>::denorm_min();
      }
    }
    const T omega = T(1) / tau;
    const T fp = T(1) - T(0.5) * omega;
    if (kEdge && !(omega > T(0.5))) {
      ++skipped;
      continue;
    }
    reference_collide(g, fx, fy, omega, fp);
    lbmhip::collide(h, fx, fy, omega, fp);
    bool finite = true;
    for (int i = 0; i < kQ; ++i) finite = finite && std::isfinite(g[i]);
    if (!finite) {
      ++skipped;
      continue;
    }
    ++compared;
    for (int i = 0; i < kQ; ++i)
      if (!same_bits(g[i], h[i])) {
        if (bad < 5)
          std::printf("%s mismatch case %ld dir %d: reference %a library %a\n", name, n, i,
                      static_cast<double>(g[i]), static_cast<double>(h[i]));
        ++bad;
        break;
      }
  }
  std::printf("%s: %ld cases compared, %ld non-finite skipped, %ld mismatches\n", name,
              compared, skipped, bad);
  return bad == 0 ? 0 : 1;
}

}  // namespace

int main() {
  int rc = run<float, false>("float collide", 2000000);
  rc |= run<double, false>("double collide", 2000000);
  rc |= run<float, true>("float collide, tau near 2 and -0 / subnormal pairs", 2000000);
  rc |= run<double, true>("double collide, tau near 2 and -0 / subnormal pairs", 2000000);
  std::printf(rc == 0 ? "ALL PASSED\n" : "FAILED\n");
  return rc;
}
