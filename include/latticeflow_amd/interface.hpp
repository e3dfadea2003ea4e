// C++ type-erased interface of the reference (solver/interface.hpp:15-191),
// implemented over the C ABI of liblatticeflow_amd.so (latticeflow_amd.h).
//
// Same namespace, class, typedef and function names and signatures as the
// reference, so code written against its interface.hpp compiles against this
// header and links against the MI355X library instead:
//
//   latticeQBP::LatticeFlowInterface   non-virtual accessors nDimensions,
//       shape, nNodes, nFillingEdges, edgeDeltas, nodeIndex, unary_E{0,1}_index,
//       binary_E{00,01,10,11}_index; virtual addEnergyPotentials, run,
//       runTime, getCut (interface.hpp:15-144)
//   latticeQBP::solver_ptr             std::shared_ptr<LatticeFlowInterface> (:147)
//   latticeQBP::BadSolverOrKernelName  (:152-168)
//   latticeQBP::getLatticeFlowInterface (1-d, 2-d, 3-d, 4-d and size_t* forms,
//       :171-183), isValidSolver, isValidKernel, kernelDimension,
//       validKernelNames, validSolverNames (:185-190)
//
// getCut() returns the canonical minimal source side (the reference returns
// its schedule-dependent flipped set; both are minimum cuts).  Failures other
// than bad names throw std::runtime_error with the library's message.
#pragma once

#include <cstddef>
#include <cstdint>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../latticeflow_amd.h"

namespace latticeQBP {
using std::string;
using std::vector;

class BadSolverOrKernelName : public std::exception {
 public:
  BadSolverOrKernelName(const string& _name) : msg(string("Bad Kernel or Solver Name: ") + _name) {}
  virtual ~BadSolverOrKernelName() throw() {}
  virtual const char* what() const throw() { return msg.c_str(); }

 private:
  string msg;
};

namespace amd_detail {
// LF_ERR_* -> the reference's exception types
inline void check(int rc) {
  if (rc >= 0) return;
  const string msg = lf_last_error();
  if (rc == LF_ERR_BAD_NAME) {
    const string pre = "Bad Kernel or Solver Name: ";
    throw BadSolverOrKernelName(msg.compare(0, pre.size(), pre) == 0 ? msg.substr(pre.size()) : msg);
  }
  if (rc == LF_ERR_OUT_OF_RANGE) throw std::out_of_range(msg);
  if (rc == LF_ERR_BAD_ARG || rc == LF_ERR_NONSUBMODULAR || rc == LF_ERR_OVERFLOW) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

inline vector<string> split_names(int (*fn)(char*, size_t)) {
  const int n = fn(nullptr, 0);
  string buf((size_t)n + 1, '\0');
  fn(&buf[0], buf.size());
  buf.resize((size_t)n);
  vector<string> out;
  size_t pos = 0;
  while (pos <= buf.size() && !buf.empty()) {
    const size_t e = buf.find('\n', pos);
    out.push_back(buf.substr(pos, e == string::npos ? string::npos : e - pos));
    if (e == string::npos) break;
    pos = e + 1;
  }
  return out;
}
}  // namespace amd_detail

class LatticeFlowInterface {
 public:
  virtual ~LatticeFlowInterface() {}

  size_t nDimensions() const { return _n_dimensions; }
  const vector<size_t>& shape() const { return _shape; }
  size_t nNodes() const { return _n_nodes; }
  size_t nFillingEdges() const { return _n_filling_edges; }
  // entry [i * nDimensions() + j]: offset of filling edge i in dimension j
  const vector<int>& edgeDeltas() const { return _edge_deltas; }

  // unary_energy_array[2 n + {0,1}] = E0, E1; binary_energy_array[n (4 k/2) +
  // 4 e + {0..3}] = E00, E01, E10, E11 (interface.hpp:46-61)
  virtual void addEnergyPotentials(long* unary_energy_array, long* binary_energy_array) = 0;

  inline size_t nodeIndex(size_t* node_coords) const {
    size_t mult = 1, node_index = 0;
    for (size_t dim = _n_dimensions; (dim--) != 0;) {
      node_index += mult * node_coords[dim];
      mult *= _shape[dim];
    }
    return node_index;
  }
  inline size_t unary_E0_index(size_t node_index) const { return 2 * node_index + 0; }
  inline size_t unary_E1_index(size_t node_index) const { return 2 * node_index + 1; }
  inline size_t binary_E00_index(size_t n, size_t e) const { return n * (4 * _n_filling_edges) + 4 * e + 0; }
  inline size_t binary_E01_index(size_t n, size_t e) const { return n * (4 * _n_filling_edges) + 4 * e + 1; }
  inline size_t binary_E10_index(size_t n, size_t e) const { return n * (4 * _n_filling_edges) + 4 * e + 2; }
  inline size_t binary_E11_index(size_t n, size_t e) const { return n * (4 * _n_filling_edges) + 4 * e + 3; }

  virtual void run() = 0;
  virtual double runTime() = 0;
  virtual vector<int> getCut() const = 0;

 protected:
  LatticeFlowInterface() {}
  void init(size_t n_dimensions, size_t n_filling_edges, vector<size_t> shape, vector<int> edge_deltas) {
    _n_dimensions = n_dimensions;
    _n_filling_edges = n_filling_edges;
    _n_nodes = 1;
    for (size_t s : shape) _n_nodes *= s;
    _shape = shape;
    _edge_deltas = edge_deltas;
  }

 private:
  size_t _n_dimensions = 0, _n_filling_edges = 0, _n_nodes = 0;
  vector<size_t> _shape;
  vector<int> _edge_deltas;
};

typedef std::shared_ptr<LatticeFlowInterface> solver_ptr;

// The library's solver behind the interface: one lf_handle.
class AmdLatticeFlow : public LatticeFlowInterface {
 public:
  AmdLatticeFlow(const string& solver, const string& kernel, const vector<size_t>& dims, int cap_bytes = 8) {
    if (!lf_is_valid_solver(solver.c_str())) throw BadSolverOrKernelName(solver);
    if (!lf_is_valid_kernel(kernel.c_str())) throw BadSolverOrKernelName(kernel);
    vector<int64_t> shape(dims.begin(), dims.end());
    amd_detail::check(lf_create(&h_, solver.c_str(), kernel.c_str(), (int)shape.size(), shape.data(), cap_bytes));
    const int ne = lf_n_filling_edges(h_);
    vector<int32_t> ed((size_t)ne * dims.size());
    amd_detail::check(lf_edge_deltas(h_, ed.data()));
    init(dims.size(), (size_t)ne, dims, vector<int>(ed.begin(), ed.end()));
  }
  ~AmdLatticeFlow() override {
    if (h_) lf_destroy(h_);
  }
  AmdLatticeFlow(const AmdLatticeFlow&) = delete;
  AmdLatticeFlow& operator=(const AmdLatticeFlow&) = delete;

  void addEnergyPotentials(long* unary, long* binary) override {
    static_assert(sizeof(long) == sizeof(int64_t), "long must be 64-bit (LP64)");
    amd_detail::check(lf_add_energy_potentials(h_, reinterpret_cast<const int64_t*>(unary),
                                               reinterpret_cast<const int64_t*>(binary)));
  }
  void run() override { amd_detail::check(lf_run(h_)); }
  double runTime() override {
    double t = 0;
    amd_detail::check(lf_run_time(h_, &t));
    return t;
  }
  vector<int> getCut() const override {
    vector<int32_t> c(nNodes());
    amd_detail::check(lf_get_cut(h_, c.data(), LF_CUT_SMIN));
    return vector<int>(c.begin(), c.end());
  }
  // additions of this library: the maximal cut and the raw max-flow value
  vector<int> getCutMax() const {
    vector<int32_t> c(nNodes());
    amd_detail::check(lf_get_cut(h_, c.data(), LF_CUT_SMAX));
    return vector<int>(c.begin(), c.end());
  }
  long flow() const {
    int64_t f = 0;
    amd_detail::check(lf_get_flow(h_, &f));
    return (long)f;
  }
  lf_handle* handle() const { return h_; }

 private:
  lf_handle* h_ = nullptr;
};

inline bool isValidSolver(const string& _name) { return lf_is_valid_solver(_name.c_str()) == 1; }
inline bool isValidKernel(const string& _name) { return lf_is_valid_kernel(_name.c_str()) == 1; }
inline int kernelDimension(const string& kernel) {
  const int d = lf_kernel_dimension(kernel.c_str());
  if (d < 0) throw BadSolverOrKernelName(kernel);
  return d;
}
inline vector<string> validKernelNames() { return amd_detail::split_names(lf_valid_kernel_names); }
inline vector<string> validSolverNames() { return amd_detail::split_names(lf_valid_solver_names); }

// Generic interface: the kernel fixes how many entries of dimensions are read
inline solver_ptr getLatticeFlowInterface(const string& solver, const string& kernel, size_t* dimensions) {
  const int nd = kernelDimension(kernel);
  return solver_ptr(new AmdLatticeFlow(solver, kernel, vector<size_t>(dimensions, dimensions + nd)));
}
inline solver_ptr getLatticeFlowInterface(const string& solver, const string& kernel, size_t nx) {
  return solver_ptr(new AmdLatticeFlow(solver, kernel, {nx}));
}
inline solver_ptr getLatticeFlowInterface(const string& solver, const string& kernel, size_t nx, size_t ny) {
  return solver_ptr(new AmdLatticeFlow(solver, kernel, {nx, ny}));
}
inline solver_ptr getLatticeFlowInterface(const string& solver, const string& kernel, size_t nx, size_t ny,
                                          size_t nz) {
  return solver_ptr(new AmdLatticeFlow(solver, kernel, {nx, ny, nz}));
}
inline solver_ptr getLatticeFlowInterface(const string& solver, const string& kernel, size_t nx1, size_t nx2,
                                          size_t nx3, size_t nx4) {
  return solver_ptr(new AmdLatticeFlow(solver, kernel, {nx1, nx2, nx3, nx4}));
}

}  // namespace latticeQBP
