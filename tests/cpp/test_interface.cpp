// Drives the C++ interface headers (include/latticeflow_amd/interface.hpp,
// solvers.hpp) the way code written against the reference's solver/interface.hpp
// and solvers.hpp does, over the library's C ABI.  Run by
// tests/test_gpu_cpp_interface.py:
//
//   test_interface names                 names, dimensions and the bad-name
//                                        exception (no device work)
//   test_interface solve DIR KERNEL SOLVER NX NY
//        reads DIR/unary.bin (int64 [n][2]) and DIR/binary.bin (int64
//        [n][k/2][4]), builds the solver with getLatticeFlowInterface, writes
//        DIR/cut_<SOLVER>.bin (int32 [n]) and DIR/meta_<SOLVER>.txt (the
//        interface's counts and edge deltas)
//   test_interface tv DIR KERNEL NX NY LAMBDA...
//        reads DIR/f.bin (double [nx][ny]); calculate2dTV<Kernel, int64_t> at
//        the first lambda -> DIR/tv.bin, and over all lambdas -> DIR/path.bin
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "latticeflow_amd/interface.hpp"
#include "latticeflow_amd/solvers.hpp"

using namespace latticeQBP;

template <typename T>
static std::vector<T> read_all(const std::string& path) {
  std::ifstream in(path, std::ios::binary | std::ios::ate);
  if (!in) throw std::runtime_error("cannot read " + path);
  const size_t bytes = (size_t)in.tellg();
  in.seekg(0);
  std::vector<T> v(bytes / sizeof(T));
  in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
  return v;
}

template <typename T>
static void write_all(const std::string& path, const T* p, size_t n) {
  std::ofstream out(path, std::ios::binary);
  out.write(reinterpret_cast<const char*>(p), (std::streamsize)(n * sizeof(T)));
}

template <class Kernel>
static int run_tv(const std::string& dir, size_t nx, size_t ny, const std::vector<double>& lams) {
  std::vector<double> f = read_all<double>(dir + "/f.bin");
  if (f.size() != nx * ny) throw std::runtime_error("f.bin has the wrong size");
  FuncMapPtr R = calculate2dTV<Kernel, int64_t>(nx, ny, f.data(), lams[0]);
  std::vector<double> r(nx * ny);
  for (size_t i = 0; i < nx; ++i)
    for (size_t j = 0; j < ny; ++j) r[i * ny + j] = R->at(i, j);
  write_all(dir + "/tv.bin", r.data(), r.size());
  RegPathPtr P = calculate2dTV<Kernel, int64_t>(nx, ny, f.data(), lams);
  std::vector<double> p(lams.size() * nx * ny);
  for (size_t l = 0; l < lams.size(); ++l)
    for (size_t i = 0; i < nx; ++i)
      for (size_t j = 0; j < ny; ++j) p[(l * nx + i) * ny + j] = P->at(l, i, j);
  write_all(dir + "/path.bin", p.data(), p.size());
  std::cout << "tv ok" << std::endl;
  return 0;
}

int main(int argc, char** argv) {
  try {
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "names") {
      for (const std::string& s : validSolverNames()) std::cout << "solver " << s << "\n";
      std::cout << "kernels " << validKernelNames().size() << "\n";
      std::cout << "dim Full2d_8 " << kernelDimension("Full2d_8") << " Star3d_26 " << kernelDimension("Star3d_26")
                << " Full4d_8 " << kernelDimension("Full4d_8") << "\n";
      std::cout << "valid " << isValidSolver("LatticeEnergyMinimizer") << isValidKernel("Full2d_8")
                << isValidKernel("Nope2d_3") << "\n";
      try {
        getLatticeFlowInterface("NoSuchSolver", "Full2d_8", 4, 4);
        std::cout << "no exception\n";
      } catch (const BadSolverOrKernelName& e) {
        std::cout << "exception " << e.what() << "\n";
      }
      return 0;
    }
    if (mode == "solve" && argc == 7) {
      const std::string dir = argv[2], kernel = argv[3], solver = argv[4];
      const size_t nx = std::strtoull(argv[5], nullptr, 10), ny = std::strtoull(argv[6], nullptr, 10);
      solver_ptr lf = getLatticeFlowInterface(solver, kernel, nx, ny);
      std::vector<long> unary = read_all<long>(dir + "/unary.bin");
      std::vector<long> binary = read_all<long>(dir + "/binary.bin");
      if (unary.size() != 2 * lf->nNodes() || binary.size() != 4 * lf->nNodes() * lf->nFillingEdges())
        throw std::runtime_error("energy files do not match the lattice");
      // fill through the interface's index helpers, as a reference caller does
      std::vector<long> u(unary.size()), b(binary.size());
      size_t coords[2];
      for (coords[0] = 0; coords[0] < nx; ++coords[0])
        for (coords[1] = 0; coords[1] < ny; ++coords[1]) {
          const size_t n = lf->nodeIndex(coords);
          u[lf->unary_E0_index(n)] = unary[2 * n];
          u[lf->unary_E1_index(n)] = unary[2 * n + 1];
          for (size_t e = 0; e < lf->nFillingEdges(); ++e) {
            const size_t base = (n * lf->nFillingEdges() + e) * 4;
            b[lf->binary_E00_index(n, e)] = binary[base + 0];
            b[lf->binary_E01_index(n, e)] = binary[base + 1];
            b[lf->binary_E10_index(n, e)] = binary[base + 2];
            b[lf->binary_E11_index(n, e)] = binary[base + 3];
          }
        }
      lf->addEnergyPotentials(u.data(), b.data());
      lf->run();
      const std::vector<int> cut = lf->getCut();
      write_all(dir + "/cut_" + solver + ".bin", cut.data(), cut.size());
      std::ofstream meta(dir + "/meta_" + solver + ".txt");
      meta << lf->nDimensions() << " " << lf->nNodes() << " " << lf->nFillingEdges();
      for (int d : lf->edgeDeltas()) meta << " " << d;
      meta << "\n" << lf->runTime() << "\n";
      std::cout << "solve ok " << solver << std::endl;
      return 0;
    }
    if (mode == "tv" && argc >= 7) {
      const std::string dir = argv[2], kernel = argv[3];
      const size_t nx = std::strtoull(argv[4], nullptr, 10), ny = std::strtoull(argv[5], nullptr, 10);
      std::vector<double> lams;
      for (int i = 6; i < argc; ++i) lams.push_back(std::atof(argv[i]));
      if (kernel == "Full2d_8") return run_tv<Full2d_8>(dir, nx, ny, lams);
      if (kernel == "Star2d_16") return run_tv<Star2d_16>(dir, nx, ny, lams);
      if (kernel == "Full2d_4") return run_tv<Full2d_4>(dir, nx, ny, lams);
      throw BadSolverOrKernelName(kernel);
    }
    std::cerr << "usage: test_interface names | solve DIR KERNEL SOLVER NX NY | tv DIR KERNEL NX NY LAMBDA...\n";
    return 2;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
}
