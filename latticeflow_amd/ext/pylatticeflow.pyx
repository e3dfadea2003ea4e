# distutils: language = c++
# cython: language_level=3, boundscheck=False, wraparound=False
"""pylatticeflow: the reference's Cython extension (extensions/python/
pylatticeflow.pyx:1-484) for the MI355X library.

Same module and class names, same arguments, same ValueError messages and the
same scaling quirks of calculate2dTV (times lambda) and calculate2dTVPath
(divided by the last lambda of the list).  The C++ side is the library's
interface.hpp / solvers.hpp (getLatticeFlowInterface, calculate2dTV
templates), which call the device library through its C ABI; run() releases
the GIL.  getCut() returns the minimal source side (the reference's flipped
set depends on its schedule; both are minimum cuts)."""
from libcpp.memory cimport shared_ptr
from libcpp.string cimport string
from libcpp.vector cimport vector

import numpy as np

cdef extern from "latticeflow_amd/interface.hpp" namespace "latticeQBP":
    cdef cppclass LatticeFlowInterface:
        size_t nDimensions()
        size_t nNodes()
        size_t nFillingEdges()
        const vector[int]& edgeDeltas()
        void addEnergyPotentials(long* unary_energy_array, long* binary_energy_array) except +
        void run() except + nogil
        double runTime() except +
        vector[int] getCut() except +

    shared_ptr[LatticeFlowInterface] getLatticeFlowInterface(const string& solver, const string& kernel,
                                                             size_t* dimensions) except +
    bint isValidSolver(const string& name)
    bint isValidKernel(const string& name)
    int kernelDimension(const string& name) except +
    vector[string] validKernelNames()
    vector[string] validSolverNames()

cdef extern from "latticeflow_amd/solvers.hpp" namespace "latticeQBP":
    cdef cppclass FuncArray "latticeQBP::LatticeArray<double, 2>":
        double at(size_t, size_t) except + nogil
    cdef cppclass FuncPathArray "latticeQBP::LatticeArray<double, 3>":
        double at(size_t, size_t, size_t) except + nogil
    shared_ptr[FuncArray] _calculate2dTV "latticeQBP::calculate2dTV<latticeQBP::Star2d_16, int64_t>" (
        size_t nx, size_t ny, double* function, double lm) except + nogil
    shared_ptr[FuncPathArray] _calculate2dTVPath "latticeQBP::calculate2dTV<latticeQBP::Star2d_16, int64_t>" (
        size_t nx, size_t ny, double* function, vector[double] lm) except + nogil


cdef list _names(vector[string] v):
    return [s.decode() for s in v]


cpdef getKernelDimension(str kernel):
    if not isValidKernel(kernel.encode()):
        raise ValueError("Kernel '%s' not a valid kernel name!  Valid kernels are: %s."
                         % (kernel, ", ".join(_names(validKernelNames()))))
    return kernelDimension(kernel.encode())


cdef class LatticeFlow(object):
    """An interface to the lattice flow solvers (pyx:82-302)."""

    cdef shared_ptr[LatticeFlowInterface] lattice_flow
    cdef size_t n_nodes
    cdef tuple _dimensions
    cdef double numerical_precision_factor
    cdef object _edges

    def __init__(self, str kernel, str solver, _dimensions, numerical_precision=1e-6):
        if not isValidSolver(solver.encode()):
            raise ValueError("'%s' not a valid solver name!  Valid solvers are: %s."
                             % (solver, ", ".join(_names(validSolverNames()))))
        cdef size_t dim = getKernelDimension(kernel)
        self._dimensions = tuple(int(d) for d in _dimensions)
        self.n_nodes = int(np.prod(self._dimensions))
        if len(self._dimensions) != dim:
            raise ValueError("Kernel '%s' is %d-dimensional; a shape of %d dimensions was provided."
                             % (kernel, dim, len(self._dimensions)))
        cdef size_t dimensions[16]
        cdef size_t i
        for i in range(dim):
            dimensions[i] = self._dimensions[i]
        self.lattice_flow = getLatticeFlowInterface(solver.encode(), kernel.encode(), &dimensions[0])
        self.numerical_precision_factor = 1.0 / numerical_precision
        cdef size_t ne = self.lattice_flow.get().nFillingEdges()
        cdef size_t nd = self.lattice_flow.get().nDimensions()
        self._edges = np.array(list(self.lattice_flow.get().edgeDeltas()), dtype=np.int32).reshape(ne, nd)

    @property
    def ndim(self):
        return self.lattice_flow.get().nDimensions()

    @property
    def shape(self):
        return self._dimensions

    @property
    def size(self):
        return self.n_nodes

    @property
    def nFillingEdges(self):
        return self.lattice_flow.get().nFillingEdges()

    @property
    def edges(self):
        return self._edges

    def addEnergy(self, unary_energy, pairwise_energy):
        """pyx:157-267: scale by 1/numerical_precision, round to long, validate, add."""
        from latticeflow_amd.pylatticeflow import _energy_arrays
        nf = self.numerical_precision_factor
        u, p = _energy_arrays(self._dimensions, self.n_nodes, self.nFillingEdges,
                              np.round(nf * np.asarray(unary_energy)), np.round(nf * np.asarray(pairwise_energy)))
        cdef long[:, ::1] unary_a = u
        cdef long[:, :, ::1] pairwise_a = p
        if self.n_nodes == 0:
            return
        self.lattice_flow.get().addEnergyPotentials(&unary_a[0, 0], &pairwise_a[0, 0, 0])

    def run(self):
        """Runs the specified solver (the GIL is released meanwhile)."""
        cdef LatticeFlowInterface* lf = self.lattice_flow.get()
        with nogil:
            lf.run()

    def runTime(self):
        """Wall seconds of the last call to run (the reference reports CPU time)."""
        return self.lattice_flow.get().runTime()

    def getCut(self):
        """The cut as an int32 array of the lattice's shape; 1 = on (source side)."""
        cdef vector[int] cut = self.lattice_flow.get().getCut()
        out = np.empty(cut.size(), dtype=np.int32)
        cdef int[::1] o = out
        cdef size_t i
        for i in range(cut.size()):
            o[i] = cut[i]
        return out.reshape(*self._dimensions)


def _runBenchmark(image_file, str kernel, str solver, bint check_result=False):
    """pyx:305-435 on this module's LatticeFlow."""
    import hashlib
    from latticeflow_amd.pylatticeflow import _benchmark_energies
    if isinstance(image_file, str):
        from matplotlib.pyplot import imread
        Xo = imread(image_file)
        if not Xo.size:
            raise IOError("Error loading image %s." % image_file)
    else:
        Xo = np.asarray(image_file)
    if Xo.ndim != 3:
        raise IOError("Error loading image; need RGB images.")
    dims = (Xo.shape[0], Xo.shape[1]) if getKernelDimension(kernel) == 2 else (Xo.shape[0], Xo.shape[1], Xo.shape[1])
    lf = LatticeFlow(kernel, solver, dims)
    print("Calculating potentials.")
    shape, unary, binary = _benchmark_energies(Xo, kernel, lf.edges)
    lf.addEnergy(unary, binary)
    lf.run()
    cut = lf.getCut().ravel()
    runtime = lf.runTime()
    print("Final result: %d / %d turned on, solution hash = %s, ran in %f seconds."
          % (cut.sum(), cut.shape[0], hashlib.md5(cut.tobytes()).hexdigest()[:8], runtime))
    if check_result:
        print("Checking solution:")
        lf2 = LatticeFlow(kernel, "BKGCEnergyMinimizer", dims)
        lf2.addEnergy(unary, binary)
        lf2.run()
        print("BKGC Solver ran in %f seconds." % lf2.runTime())
        cut2 = lf2.getCut().ravel()
        if (cut != cut2).sum() != 0:
            print("Solution different: %d falsely 0, %d falsely 1"
                  % (((cut == 0) & (cut2 == 1)).sum(), ((cut == 1) & (cut2 == 0)).sum()))
        else:
            print("Solutions exactly match.")
    return lf.runTime()


def calculate2dTV(Xo, double flt):
    """pyx:437-454: calculate2dTV<Star2d_16, int64_t>, times lambda."""
    X = np.ascontiguousarray(Xo, dtype=np.float64)
    cdef double[:, ::1] Xv = X
    cdef size_t nx = X.shape[0], ny = X.shape[1]
    cdef shared_ptr[FuncArray] Rv
    with nogil:
        Rv = _calculate2dTV(nx, ny, &Xv[0, 0], flt)
    R = np.empty((nx, ny))
    cdef double[:, ::1] Rw = R
    cdef size_t xi, yi
    for xi in range(nx):
        for yi in range(ny):
            Rw[xi, yi] = Rv.get().at(xi, yi) * flt
    return R


def calculate2dTVPath(Xo, lma):
    """pyx:457-484: the path at the lambdas; like the reference, every slice is
    divided by the last lambda of the list (its loop index is stale there)."""
    X = np.ascontiguousarray(Xo, dtype=np.float64)
    lam = np.ascontiguousarray(lma, dtype=np.float64).reshape(-1)
    cdef double[:, ::1] Xv = X
    cdef size_t nx = X.shape[0], ny = X.shape[1]
    cdef size_t n_lm = lam.shape[0]
    cdef vector[double] lm_values = vector[double](n_lm)
    cdef size_t i
    for i in range(n_lm):
        lm_values[i] = lam[i]
    cdef shared_ptr[FuncPathArray] Rv
    with nogil:
        Rv = _calculate2dTVPath(nx, ny, &Xv[0, 0], lm_values)
    R = np.empty((n_lm, nx, ny))
    if n_lm == 0:
        return R
    cdef double[:, :, ::1] Rw = R
    cdef double last = lam[n_lm - 1]
    cdef size_t li, xi, yi
    for li in range(n_lm):
        for xi in range(nx):
            for yi in range(ny):
                Rw[li, xi, yi] = Rv.get().at(li, xi, yi) / last
    return R
