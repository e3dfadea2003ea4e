// lf_mincut_wide.hip -- device min-cut engines for the wider neighbourhoods:
// 1-D and 2-D kernels of radius 2 to 4 with up to 28 directions
// (Full1d_4/6/8, Full2d_20/24, Star2d_24, Full2d_28; Full2d_12 and Star2d_16
// are built by lf_mincut.hip, 32 directions and more by lf_mincut_wide32.hip).
// The kernels are those of lf_mincut.hip, instantiated with push tiles small
// enough for k residual planes in LDS (k x 8 B x 512 nodes = 112 KB for
// Full2d_28 in int64); a translation unit of its own, so that the two compile
// in parallel.
#define LF_MINCUT_WIDE_TU
#include "lf_mincut.hip"

namespace lfamd {

std::unique_ptr<MinCutEngine> make_wide_engine(const LatticeGeom& g, int cap_bytes) {
  const int rz = max_abs(g.dz), ry = max_abs(g.dy), rx = max_abs(g.dx);
  if (g.ndim == 3 || rz != 0) return make_wide32_engine(g, cap_bytes);
  const int r = std::max(ry, rx);
  if (r == 2) {
    using G0 = Geo<1, 8, 64, 0, 2, 2>;
    using G1 = Geo<1, 8, 32, 0, 2, 2>;
    using G2 = Geo<1, 16, 32, 0, 2, 2>;
    using B = BGeo<1, 64, 64, 0, 2, 2>;
    if (g.k == 4) return make_k<4, G0, G1, G2, B>(g, cap_bytes);
    if (g.k == 20) return make_k<20, G0, G1, G2, B>(g, cap_bytes);
    if (g.k == 24) return make_k<24, G0, G1, G2, B>(g, cap_bytes);
  }
  if (r == 3) {
    using G0 = Geo<1, 8, 64, 0, 3, 3>;
    using G1 = Geo<1, 8, 32, 0, 3, 3>;
    using G2 = Geo<1, 16, 32, 0, 3, 3>;
    using B = BGeo<1, 64, 64, 0, 3, 3>;
    if (g.k == 6) return make_k<6, G0, G1, G2, B>(g, cap_bytes);
    if (g.k == 24) return make_k<24, G0, G1, G2, B>(g, cap_bytes);
    if (g.k == 28) return make_k<28, G0, G1, G2, B>(g, cap_bytes);
  }
  if (r == 4) {
    using G0 = Geo<1, 8, 64, 0, 4, 4>;
    using G1 = Geo<1, 8, 32, 0, 4, 4>;
    using G2 = Geo<1, 16, 32, 0, 4, 4>;
    using B = BGeo<1, 64, 64, 0, 4, 4>;
    if (g.k == 8) return make_k<8, G0, G1, G2, B>(g, cap_bytes);
  }
  return make_wide32_engine(g, cap_bytes);
}

}  // namespace lfamd
