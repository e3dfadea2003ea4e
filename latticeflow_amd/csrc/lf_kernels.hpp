// Kernel (neighbourhood) family and lattice edge order.
//
// Restates, in C++, what the reference generates at configure time:
//   * the Full / Star offset sets and their names "%(type)s%(dim)dd_%(n)d"
//     (install/code_setup.py:40, :199-256), over the radii of
//     install/code_setup.py:75-78 (2D up to radius 4, 3D up to radius 2, and
//     the 4D kernels of radius 1 that --4d adds, :141-153, wscript:64-67);
//   * the Cauchy-Crofton geocut weights of 2D kernels (install/code_setup.py:163-197);
//   * the lattice edge order: directions sorted by their linear jump over the
//     padded strides, positive half ascending, entry i + k/2 = -entry i
//     (solver/lattices/kernellattice.hpp:61-96, padding solver/lattices/lattice.hpp:102-111).
//
// The generator evaluates `0 < |x|^2 <= radius**2` in IEEE doubles, and
// sqrt(3)**2 < 3 and sqrt(13)**2 < 13 there, so it never produces a
// 26-connected "Full3d_26" or a 44-edge "Full2d_44": the 26-connected 3D
// kernel is "Star3d_26".  The same arithmetic is used here, and "Full3d_26"
// is accepted as an alias of "Star3d_26".
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <set>
#include <string>
#include <vector>

namespace lfamd {

struct KernelDef {
  std::string name;
  int ndim = 0;
  std::vector<std::vector<int>> edges;  // generator order (lexicographic)
  bool geocut = false;
  std::vector<double> geocut_w;         // per edge, generator order
};

namespace detail {

inline std::vector<double> radius_intervals(int dim, int max_radius) {
  std::set<double> vals;
  std::vector<int> x(dim, 0);
  for (;;) {
    long s = 0;
    for (int v : x) s += (long)v * v;
    vals.insert(std::sqrt((double)s));
    int d = 0;
    while (d < dim && ++x[d] > max_radius) { x[d] = 0; ++d; }
    if (d == dim) break;
  }
  std::vector<double> out;
  for (double d : vals)
    if (d > 0 && d <= max_radius + 1e-8) out.push_back(d);
  return out;
}

inline std::vector<std::vector<int>> full_edges(int dim, double radius) {
  int lo = (int)std::floor(-radius), hi = (int)std::ceil(radius);
  std::vector<std::vector<int>> out;
  std::vector<int> x(dim, lo);
  const double r2 = radius * radius;
  for (;;) {
    long s = 0;
    for (int v : x) s += (long)v * v;
    if (s > 0 && (double)s <= r2) out.push_back(x);
    int d = dim - 1;
    while (d >= 0 && ++x[d] > hi) { x[d] = lo; --d; }
    if (d < 0) break;
  }
  std::sort(out.begin(), out.end());
  return out;
}

inline std::vector<std::vector<int>> star_edges(int dim, double radius) {
  auto full = full_edges(dim, radius);
  std::vector<std::vector<int>> by_len = full;
  std::stable_sort(by_len.begin(), by_len.end(), [](const std::vector<int>& a, const std::vector<int>& b) {
    long sa = 0, sb = 0;
    for (int v : a) sa += (long)v * v;
    for (int v : b) sb += (long)v * v;
    return sa < sb;
  });
  std::set<std::vector<double>> seen;
  std::set<std::vector<int>> removed;
  for (const auto& e : by_len) {
    std::vector<double> ang;
    for (int d = 1; d < dim; ++d) ang.push_back(std::atan2((double)e[d], (double)e[0]));
    if (seen.count(ang)) removed.insert(e);
    else seen.insert(ang);
  }
  std::vector<std::vector<int>> out;
  for (const auto& e : full)
    if (!removed.count(e)) out.push_back(e);
  return out;
}

inline void geocut(KernelDef& k) {
  const size_t n = k.edges.size();
  k.geocut = false;
  k.geocut_w.assign(n, 1.0);
  if (k.ndim != 2 || n < 4) return;
  std::vector<std::pair<double, size_t>> info;
  std::set<double> angset;
  for (size_t i = 0; i < n; ++i) {
    double a = std::fmod(std::atan2((double)k.edges[i][1], (double)k.edges[i][0]), 2 * M_PI);
    if (a < 0) a += 2 * M_PI;  // Python's % is a floor modulus
    info.push_back({a, i});
    angset.insert(a);
  }
  if (angset.size() != n) return;
  std::sort(info.begin(), info.end(), [&](const std::pair<double, size_t>& a, const std::pair<double, size_t>& b) {
    if (a.first != b.first) return a.first < b.first;
    return k.edges[a.second] < k.edges[b.second];
  });
  for (size_t i = 0; i < n; ++i) {
    double a1 = info[(i + n - 1) % n].first, a2 = info[(i + 1) % n].first;
    if (a2 < a1) a2 += 2 * M_PI;
    double w = (a2 - a1) / 2;
    const auto& e = k.edges[info[i].second];
    k.geocut_w[info[i].second] = w / (2.0 * std::sqrt((double)(e[0] * e[0] + e[1] * e[1])));
  }
  k.geocut = true;
}

}  // namespace detail

// All registered kernels, keyed by name (aliases included).
inline const std::map<std::string, KernelDef>& kernel_registry() {
  static const std::map<std::string, KernelDef> reg = [] {
    std::map<std::string, KernelDef> m;
    auto add = [&](const std::string& type, int dim, std::vector<std::vector<int>> e) {
      KernelDef k;
      k.ndim = dim;
      k.edges = std::move(e);
      k.name = type + std::to_string(dim) + "d_" + std::to_string(k.edges.size());
      if (m.count(k.name)) return;
      detail::geocut(k);
      m[k.name] = k;
    };
    for (int r = 1; r <= 4; ++r) add("Full", 1, detail::full_edges(1, r));
    // 2D up to radius 4, 3D up to radius 2, 4D radius 1 (the --4d default,
    // wscript:64-67)
    for (int dim = 2; dim <= 4; ++dim) {
      int maxr = dim == 2 ? 4 : (dim == 3 ? 2 : 1);
      for (double r : detail::radius_intervals(dim, maxr)) {
        add("Full", dim, detail::full_edges(dim, r));
        add("Star", dim, detail::star_edges(dim, r));
      }
    }
    KernelDef alias = m.at("Star3d_26");
    alias.name = "Full3d_26";
    m["Full3d_26"] = alias;
    return m;
  }();
  return reg;
}

inline const KernelDef* find_kernel(const std::string& name) {
  const auto& r = kernel_registry();
  auto it = r.find(name);
  return it == r.end() ? nullptr : &it->second;
}

// Directions in lattice edge order for a given unpadded shape.
// offs[i*nd + d]; weights[i]; entry i + k/2 is the reverse of entry i.
inline bool lattice_edge_order(const KernelDef& k, const std::vector<int64_t>& shape,
                               std::vector<int>& offs, std::vector<double>& weights) {
  const int nd = k.ndim;
  const size_t ks = k.edges.size();
  if ((int)shape.size() != nd) return false;
  std::vector<int64_t> eb(nd, 0), pd(nd), st(nd);
  for (int d = 0; d < nd; ++d)
    for (const auto& e : k.edges) eb[d] = std::max<int64_t>(eb[d], std::abs(e[d]));
  pd[0] = shape[0] + 4 * eb[0];
  for (int d = 1; d < nd; ++d) pd[d] = shape[d] + eb[d];
  st[nd - 1] = 1;
  for (int d = nd - 2; d >= 0; --d) st[d] = st[d + 1] * pd[d + 1];
  std::vector<int64_t> jumps(ks);
  for (size_t i = 0; i < ks; ++i) {
    int64_t j = 0;
    for (int d = 0; d < nd; ++d) j += k.edges[i][d] * st[d];
    jumps[i] = j;
  }
  std::vector<size_t> idx(ks);
  for (size_t i = 0; i < ks; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return jumps[a] > jumps[b]; });
  std::reverse(idx.begin(), idx.begin() + ks / 2);
  offs.assign(ks * nd, 0);
  weights.assign(ks, 1.0);
  for (size_t i = 0; i < ks; ++i) {
    for (int d = 0; d < nd; ++d) offs[i * nd + d] = k.edges[idx[i]][d];
    weights[i] = k.geocut_w[idx[i]];
  }
  // symmetry: entry i + k/2 must be -entry i
  for (size_t i = 0; i < ks / 2; ++i)
    for (int d = 0; d < nd; ++d)
      if (offs[(i + ks / 2) * nd + d] != -offs[i * nd + d]) return false;
  return true;
}

}  // namespace lfamd
