#include "host_common.hpp"

#include <cstdio>
#include <cstdlib>

namespace lbmhip {

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

const char* last_error() { return g_last_error.c_str(); }

Links build_links(int nx, int ny, const uint8_t* solid, bool open_x, bool wrap_y) {
  Links L;
  const size_t N = static_cast<size_t>(nx) * ny;
  L.code.assign(N, 0);
  L.nseg = (nx + 63) / 64;
  L.seg.assign(static_cast<size_t>(ny) * L.nseg, 0);
  auto is_solid = [&](int x, int y) { return solid[static_cast<size_t>(y) * nx + x] != 0; };
  for (int y = 0; y < ny; ++y) {
    const bool ghost = !wrap_y && (y == 0 || y == ny - 1);
    for (int x = 0; x < nx; ++x) {
      const size_t c = static_cast<size_t>(y) * nx + x;
      if (is_solid(x, y)) {
        L.code[c] = kLinkSolid;
        continue;
      }
      if (ghost) continue;  // never updated
      uint32_t code = 0;
      for (int i = 1; i < kQ; ++i) {
        if (open_x && x == 0 && kCx[i] == 1) continue;        // inlet
        if (open_x && x == nx - 1 && kCx[i] == -1) continue;  // outlet
        const int xs = (x - kCx[i] + nx) % nx;
        const int ys = wrap_y ? (y - kCy[i] + ny) % ny : y - kCy[i];
        if (is_solid(xs, ys)) code |= 1u << i;
      }
      if (open_x && x == 0) code |= kLinkInlet;
      if (open_x && x == nx - 1) {
        code |= kLinkOutlet;
        for (int i : {3, 6, 7}) {
          const int ys = wrap_y ? (y - kCy[i] + ny) % ny : y - kCy[i];
          if (is_solid(nx - 1, ys)) code |= outlet_bb_bit(i);
        }
      }
      L.code[c] = static_cast<uint16_t>(code);
    }
  }
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x)
      if (L.code[static_cast<size_t>(y) * nx + x]) L.seg[static_cast<size_t>(y) * L.nseg + x / 64] = 1;
  return L;
}

template <class T>
void preimage(int nx, int ny, const uint8_t* solid, bool open_x, const T* S, T* R) {
  const size_t N = static_cast<size_t>(nx) * ny;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const size_t c = static_cast<size_t>(y) * nx + x;
      for (int i = 0; i < kQ; ++i) {
        T v = T(0);
        const int xr = x + kCx[i];
        if (!solid[c] && !(open_x && (xr < 0 || xr >= nx))) {
          const int xn = (xr + nx) % nx, yn = (y + kCy[i] + ny) % ny;
          const size_t cn = static_cast<size_t>(yn) * nx + xn;
          v = solid[cn] ? S[c * kQ + kOpp[i]] : S[cn * kQ + i];
        }
        R[i * N + c] = v;
      }
    }
}
template void preimage<float>(int, int, const uint8_t*, bool, const float*, float*);
template void preimage<double>(int, int, const uint8_t*, bool, const double*, double*);

template <class T>
void preimage_uniform(int nx, int ny, const uint8_t* solid, bool open_x, const T* eq, T* R) {
  const size_t N = static_cast<size_t>(nx) * ny;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const size_t c = static_cast<size_t>(y) * nx + x;
      for (int i = 0; i < kQ; ++i) {
        T v = T(0);
        if (!solid[c]) {
          const bool leaves = open_x && (x + kCx[i] < 0 || x + kCx[i] >= nx);
          const int xn = (x + kCx[i] + nx) % nx, yn = (y + kCy[i] + ny) % ny;
          v = (!leaves && solid[static_cast<size_t>(yn) * nx + xn]) ? eq[kOpp[i]] : eq[i];
        }
        R[i * N + c] = v;
      }
    }
}
template void preimage_uniform<float>(int, int, const uint8_t*, bool, const float*, float*);
template void preimage_uniform<double>(int, int, const uint8_t*, bool, const double*, double*);

int check_block(int bx, int rows) {
  if (bx < 64 || bx > 1024 || bx % 64 != 0)
    return fail(LBM_EINVAL, "block x extent must be a multiple of the 64-wide wavefront, 64..1024");
  if (rows < 1 || rows > 4096) return fail(LBM_EINVAL, "rows per thread must be in 1..4096");
  return LBM_OK;
}

int parse_block_env(int* bx, int* by) {
  const char* env = std::getenv("LBM_BLOCK");
  if (!env) return LBM_OK;
  int x = 0, y = 0;
  if (std::sscanf(env, "%dx%d", &x, &y) != 2)
    return fail(LBM_EINVAL, "LBM_BLOCK must look like 256x8 (block x extent x rows per thread)");
  LBM_TRY(check_block(x, y));
  *bx = x;
  *by = y;
  return LBM_OK;
}

}  // namespace lbmhip
