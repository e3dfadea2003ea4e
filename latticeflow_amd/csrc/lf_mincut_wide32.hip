// lf_mincut_wide32.hip -- device min-cut engines for the neighbourhoods of 32
// and 36 directions of radius 3: Star2d_32 and Full2d_36 (Full2d_48 and
// Full3d_32 are built by lf_mincut_wide48.hip).  The kernels are those of
// lf_mincut.hip with 32/64-bit arc masks and dirty words, on push tiles of 256
// nodes; every tile is at least as deep as the radius, so a push never skips
// a tile.
#define LF_MINCUT_WIDE_TU
#include "lf_mincut.hip"

namespace lfamd {

std::unique_ptr<MinCutEngine> make_wide32_engine(const LatticeGeom& g, int cap_bytes) {
  const int rz = max_abs(g.dz), ry = max_abs(g.dy), rx = max_abs(g.dx);
  if (g.ndim != 3 && rz == 0 && std::max(ry, rx) == 3) {
    using G0 = Geo<1, 8, 32, 0, 3, 3>;
    using G1 = Geo<1, 4, 64, 0, 3, 3>;
    using G2 = Geo<1, 16, 16, 0, 3, 3>;
    using B = BGeo<1, 64, 64, 0, 3, 3>;
    if (g.k == 32) return make_k<32, G0, G1, G2, B>(g, cap_bytes);
    if (g.k == 36) return make_k<36, G0, G1, G2, B>(g, cap_bytes);
  }
  return make_wide48_engine(g, cap_bytes);
}

}  // namespace lfamd
