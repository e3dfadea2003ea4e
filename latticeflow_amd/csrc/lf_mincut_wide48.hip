// lf_mincut_wide48.hip -- device min-cut engines for Full2d_48 (radius 4) and
// Full3d_32 (3-D, radius 2): the kernels of lf_mincut.hip with 64/32-bit arc
// masks on push tiles of 256 nodes (k residual planes of 8 B x 256 nodes =
// 96 KB of LDS for Full2d_48 in int64), each at least as deep as the radius.
#define LF_MINCUT_WIDE_TU
#include "lf_mincut.hip"

namespace lfamd {

std::unique_ptr<MinCutEngine> make_wide48_engine(const LatticeGeom& g, int cap_bytes) {
  const int rz = max_abs(g.dz), ry = max_abs(g.dy), rx = max_abs(g.dx);
  const int r = std::max(ry, rx);
  if (g.ndim == 3 && rz <= 2 && r <= 2 && g.k == 32)
    return make_k<32, Geo<4, 4, 16, 2, 2, 2>, Geo<2, 8, 16, 2, 2, 2>, Geo<4, 8, 8, 2, 2, 2>, BGeo<16, 16, 16, 2, 2, 2>>(
        g, cap_bytes);
  if (g.ndim != 3 && rz == 0 && r == 4 && g.k == 48)
    return make_k<48, Geo<1, 8, 32, 0, 4, 4>, Geo<1, 4, 64, 0, 4, 4>, Geo<1, 16, 16, 0, 4, 4>, BGeo<1, 64, 64, 0, 4, 4>>(
        g, cap_bytes);
  return nullptr;
}

}  // namespace lfamd
