// The per-tile update shared by the single-domain step (step.hip) and the batched persistent
// kernel (batch.hip): both call exactly this code, so every domain of a batch computes the same
// bits as the single-domain step.
#pragma once

#include "device_common.hpp"

namespace lbmhip {

// One domain as the update sees it. Plane i of the state starts at g + i * N.
template <class T>
struct DomainView {
  const T* g;
  T* gn;
  const uint16_t* links;
  const uint8_t* segflags;
  const T* force;  // per-cell field (fx plane, fy plane) when kField
  size_t N;
  int nx, ny, nseg;
  int wrap_y;
  T omega, fp, fx, fy;
};

// Updates rows [y0, y1) of column x (one thread per column; every lane of a wave shares the
// row, so row offsets stay in SGPRs). Lanes with x >= nx compute on column nx - 1 and store
// nothing. kStoreAux: cache policy of the nine stores (kAuxNone or kAuxNt, streaming).
template <class T, bool kOpen, bool kField, int kStoreAux = kAuxNone, class PS = Planes>
__device__ __forceinline__ void update_column(const DomainView<T>& d, int x, int y0, int y1,
                                              const T* eq_in) {
  const int nx = d.nx;
  // A wave whose first lane is already past the last column has nothing to do (and its
  // segment index would lie outside the row's flags).
  if (__builtin_amdgcn_readfirstlane(x) >= nx) return;
  const bool in_x = x < nx;
  const unsigned xc = in_x ? x : nx - 1;
  constexpr unsigned B = sizeof(T);
  const Cols cols{(xc == 0 ? nx - 1 : xc - 1) * B, xc * B,
                  (xc + 1 == static_cast<unsigned>(nx) ? 0 : xc + 1) * B};
  const int seg = __builtin_amdgcn_readfirstlane(x) >> 6;  // wave-uniform 64-column segment
  PS g, gn;
  g.init(d.g, d.N);
  gn.init(d.gn, d.N);
  // Force planes (read only when kField; otherwise the resources are never used).
  const Rsrc fxr = make_rsrc(d.force), fyr = make_rsrc(d.force + (kField ? d.N : 0));

  for (int y = y0; y < y1; ++y) {
    int ym = y - 1, yp = y + 1;
    if (d.wrap_y) {
      ym = ym < 0 ? d.ny - 1 : ym;
      yp = yp >= d.ny ? 0 : yp;
    }
    const Rows rows{static_cast<unsigned>(ym * nx) * B, static_cast<unsigned>(y * nx) * B,
                    static_cast<unsigned>(yp * nx) * B};
    T v[kQ];
    pull_streamed<T>(g, rows, cols, v);
    bool active = in_x;
    if (d.segflags[static_cast<size_t>(y) * d.nseg + seg]) {  // wave-uniform branch
      const uint32_t link = in_x ? d.links[static_cast<size_t>(y) * nx + xc] : kLinkSolid;
      active = !(link & kLinkSolid);
      if (active) apply_links<T, kOpen>(link, g, rows, cols, eq_in, v);
    }
    if (active) {
      T fx = d.fx, fy = d.fy;
      if constexpr (kField) {
        fx = Stored<T>::load(fxr, cols.c, rows.c);
        fy = Stored<T>::load(fyr, cols.c, rows.c);
      }
      collide(v, fx, fy, d.omega, d.fp);
#pragma unroll
      for (int i = 0; i < kQ; ++i) gn.template store<T, kStoreAux>(i, cols.c, rows.c, v[i]);
    }
  }
}

}  // namespace lbmhip
