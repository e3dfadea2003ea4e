// lf_mincut_canon.hip -- device min-cut engines of the radius-1 neighbourhoods
// that carry most of the work (2-D 4/8 and 3-D 6/26 directions), with the
// direction table compiled into the push, delivery and exact-pass kernels
// (CDir: every neighbour offset, LDS address and tile-border test is a
// constant).  Used when the lattice's edge order is the canonical one, which
// is the case for every lattice whose extents exceed the kernel radius; the
// engines of lf_mincut.hip read the table at run time otherwise.  A
// translation unit of its own, so that it compiles in parallel.
#define LF_MINCUT_WIDE_TU
#include "lf_mincut.hip"

namespace lfamd {

std::unique_ptr<MinCutEngine> make_canon_engine(const LatticeGeom& g, int cap_bytes) {
  const int rz = max_abs(g.dz), ry = max_abs(g.dy), rx = max_abs(g.dx);
  if (rz > 1 || ry > 1 || rx > 1) return nullptr;
  const int* dz = g.dz.data();
  const int* dy = g.dy.data();
  const int* dx = g.dx.data();
  if (g.ndim == 2) {
    if (CDir<8, 2>::matches(dz, dy, dx, g.k)) return make_tiled<8, Tiles2d1, CDir<8, 2>>(g, cap_bytes);
    if (CDir<4, 2>::matches(dz, dy, dx, g.k)) return make_tiled<4, Tiles2d1, CDir<4, 2>>(g, cap_bytes);
  }
  if (g.ndim == 3) {
    if (CDir<26, 3>::matches(dz, dy, dx, g.k)) return make_tiled<26, Tiles3d26, CDir<26, 3>>(g, cap_bytes);
    if (CDir<6, 3>::matches(dz, dy, dx, g.k)) return make_tiled<6, Tiles3d6, CDir<6, 3>>(g, cap_bytes);
  }
  return nullptr;
}

}  // namespace lfamd
