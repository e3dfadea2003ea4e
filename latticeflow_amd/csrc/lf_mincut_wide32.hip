// lf_mincut_wide32.hip -- device min-cut engines for the neighbourhoods of 32
// to 48 directions: Star2d_32 and Full2d_36 (radius 3), Full2d_48 (radius 4)
// and Full3d_32 (3-D, radius 2).  The kernels are those of lf_mincut.hip with
// 64-bit arc masks and dirty words, on push tiles of 256 nodes (k residual
// planes of 8 B x 256 nodes = 96 KB of LDS for Full2d_48 in int64); every tile
// is at least as deep as the radius, so a push never skips a tile.
#define LF_MINCUT_WIDE_TU
#include "lf_mincut.hip"

namespace lfamd {

std::unique_ptr<MinCutEngine> make_wide32_engine(const LatticeGeom& g, int cap_bytes) {
  const int rz = max_abs(g.dz), ry = max_abs(g.dy), rx = max_abs(g.dx);
  const int r = std::max(ry, rx);
  if (g.ndim == 3) {
    if (rz <= 2 && r <= 2 && g.k == 32)
      return make_k<32, Geo<4, 4, 16, 2, 2, 2>, Geo<2, 8, 16, 2, 2, 2>, Geo<4, 8, 8, 2, 2, 2>,
                    BGeo<16, 16, 16, 2, 2, 2>>(g, cap_bytes);
    return nullptr;
  }
  if (rz != 0) return nullptr;
  if (r == 3) {
    using G0 = Geo<1, 8, 32, 0, 3, 3>;
    using G1 = Geo<1, 4, 64, 0, 3, 3>;
    using G2 = Geo<1, 16, 16, 0, 3, 3>;
    using B = BGeo<1, 64, 64, 0, 3, 3>;
    if (g.k == 32) return make_k<32, G0, G1, G2, B>(g, cap_bytes);
    if (g.k == 36) return make_k<36, G0, G1, G2, B>(g, cap_bytes);
  }
  if (r == 4) {
    using G0 = Geo<1, 8, 32, 0, 4, 4>;
    using G1 = Geo<1, 4, 64, 0, 4, 4>;
    using G2 = Geo<1, 16, 16, 0, 4, 4>;
    using B = BGeo<1, 64, 64, 0, 4, 4>;
    if (g.k == 48) return make_k<48, G0, G1, G2, B>(g, cap_bytes);
  }
  return nullptr;
}

}  // namespace lfamd
