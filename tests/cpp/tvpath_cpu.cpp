// CPU driver of the TV path's event loop (latticeflow_amd/csrc/lf_tvpath_core.cpp)
// for tests: every level-set check runs on the host, and the level sets at
// s = 1 come from the test (the oracle's divide and conquer) instead of the
// device.  Two commands, over binary files in DIR:
//
//   tvpath_cpu net  DIR KERNEL NX NY LMAX     f.bin -> red.bin alpha.bin offs.bin
//       the TV network at LMAX as the library builds it (red[n], alpha[n*k]
//       node-major, offsets[k*2])
//   tvpath_cpu path DIR KERNEL NX NY LMAX NL  f.bin reg.bin r.bin lams.bin -> R.bin stats.txt
//       the path from the level sets reg (int64 per node) with values r at
//       s = 1, evaluated at the NL lambdas
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "../../latticeflow_amd/csrc/lf_error.hpp"
#include "../../latticeflow_amd/csrc/lf_kernels.hpp"
#include "../../latticeflow_amd/csrc/lf_tv.hpp"
#include "../../latticeflow_amd/csrc/lf_tvpath_core.hpp"

using namespace lfamd;

template <typename T>
static std::vector<T> load(const std::string& path, size_t n) {
  std::vector<T> v(n);
  std::ifstream f(path, std::ios::binary);
  if (!f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(sizeof(T) * n))) {
    std::fprintf(stderr, "cannot read %s\n", path.c_str());
    std::exit(2);
  }
  return v;
}

template <typename T>
static void save(const std::string& path, const std::vector<T>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(sizeof(T) * v.size()));
}

int main(int argc, char** argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: tvpath_cpu net|path DIR KERNEL NX NY LMAX [NL]\n");
    return 2;
  }
  const std::string cmd = argv[1], dir = argv[2], kernel = argv[3];
  const int64_t nx = std::atoll(argv[4]), ny = std::atoll(argv[5]), n = nx * ny;
  const double lmax = std::max(std::atof(argv[6]), 1e-8);
  const KernelDef* kd = find_kernel(kernel);
  if (!kd || kd->ndim != 2 || !kd->geocut) {
    std::fprintf(stderr, "bad kernel\n");
    return 2;
  }
  const int k = (int)kd->edges.size();
  std::vector<int> offs;
  std::vector<double> w;
  lattice_edge_order(*kd, {nx, ny}, offs, w);
  const std::vector<double> f = load<double>(dir + "/f.bin", (size_t)n);
  const TvBase b = tv_base(n, f.data());
  std::vector<int64_t> acap((size_t)k, 0);
  for (int e = 0; e < k / 2; ++e) acap[e] = acap[e + k / 2] = 2 * tv_pwf(b, lmax, w[e]);
  std::vector<int32_t> nbr((size_t)n * k, -1);
  for (int64_t v = 0; v < n; ++v) {
    const int64_t y = v / ny, x = v % ny;
    for (int i = 0; i < k; ++i) {
      const int64_t yy = y + offs[2 * i], xx = x + offs[2 * i + 1];
      if (yy >= 0 && yy < nx && xx >= 0 && xx < ny) nbr[(size_t)v * k + i] = (int32_t)(yy * ny + xx);
    }
  }
  if (cmd == "net") {
    std::vector<int64_t> red((size_t)n), alpha((size_t)n * k, 0);
    for (int64_t v = 0; v < n; ++v) {
      red[v] = 2 * b.uf[v];
      for (int i = 0; i < k; ++i)
        if (nbr[(size_t)v * k + i] >= 0) alpha[(size_t)v * k + i] = acap[i];
    }
    std::vector<int32_t> o(offs.begin(), offs.end());
    save(dir + "/red.bin", red);
    save(dir + "/alpha.bin", alpha);
    save(dir + "/offs.bin", o);
    return 0;
  }
  if (cmd != "path" || argc < 8) return 2;
  const int nl = std::atoi(argv[7]);
  TvPathProblem P;
  P.n = n;
  P.k = k;
  P.nbr = nbr;
  P.acap = acap;
  P.fv.resize((size_t)n);
  for (int64_t v = 0; v < n; ++v) P.fv[v] = 2 * b.uf[v];
  const std::vector<int64_t> reg = load<int64_t>(dir + "/reg.bin", (size_t)n);
  P.reg.assign(reg.begin(), reg.end());
  P.r = load<int64_t>(dir + "/r.bin", (size_t)n);
  const std::vector<double> lams = load<double>(dir + "/lams.bin", (size_t)nl);
  TvPathStats st;
  try {
    TvPathCore core(std::move(P), nullptr, 0, &st);
    std::vector<double> R((size_t)nl * n), rp((size_t)n);
    for (int l = 0; l < nl; ++l) {
      const double s = std::min(std::max(lams[l], 0.0) / lmax, 1.0);
      core.values(s, rp.data());
      tv_output_path(b, rp.data(), n, R.data() + (size_t)l * n);
    }
    save(dir + "/R.bin", R);
  } catch (const LfError& e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.msg.c_str());
    return 1;
  }
  std::ofstream so(dir + "/stats.txt");
  so << "initial_regions " << st.initial_regions << "\nsegments " << st.segments << "\njoins " << st.joins
     << "\nsplits " << st.splits << "\nchecks " << st.checks << "\ndinkelbach_steps " << st.dinkelbach_steps
     << "\nhost_check_seconds " << st.host_check_seconds << "\n";
  return 0;
}
