"""Python API of the reference's ``pylatticeflow`` extension, on the MI355X library.

Mirrors ``extensions/python/pylatticeflow.pyx`` of hoytak/latticeflow:

* ``getKernelDimension(kernel)``                      -- pyx:71-80
* ``LatticeFlow(kernel, solver, shape, numerical_precision=1e-6)`` -- pyx:82-302
  with ``ndim / shape / size / nFillingEdges / edges``, ``addEnergy``,
  ``run``, ``runTime``, ``getCut``.

Same argument meaning and the same ``ValueError`` messages for malformed
energies.  Additions, all keyword-only or new methods: ``cap_bytes`` (device
integer width; 8 = the reference's ``long``), ``getCut(which=...)`` to choose
S_min (default) or S_max, ``flow()`` (raw max-flow value), ``stats()`` and
``setTuning``.  Every call goes through the C ABI to the HIP library; there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib


def _torch_stream(device=None):
    """Handle of torch's current HIP stream (0, the null stream, without torch)."""
    try:
        import torch
    except ImportError:
        return 0
    if not torch.cuda.is_available():
        return 0
    return int(torch.cuda.current_stream(device).cuda_stream)


def _names(fn):
    n = fn(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    fn(buf, n + 1)
    return [s for s in buf.value.decode().split("\n") if s]


def validKernelNames():
    return _names(lib().lf_valid_kernel_names)


def validSolverNames():
    return _names(lib().lf_valid_solver_names)


def getKernelDimension(kernel: str) -> int:
    """pyx:71-80: ValueError listing the valid kernels on a bad name."""
    if not lib().lf_is_valid_kernel(kernel.encode()):
        raise ValueError("Kernel '%s' not a valid kernel name!  Valid kernels are: %s."
                         % (kernel, ", ".join(validKernelNames())))
    return lib().lf_kernel_dimension(kernel.encode())


def _energy_arrays(dimensions, n_nodes, n_filling_edges, unary_energy, pairwise_energy):
    """pyx:157-267: validate integral energies of a lattice of ``dimensions``
    and return them as contiguous int64 arrays (n, 2) and (n, k/2, 4), raising
    the reference's ValueError messages on malformed shapes."""
    unary_energy_error = (
        "Unary energy must be either a 2d array with first dimension "
        "matching the number of nodes and second dimension containing E0 and E1, "
        "or (#dimensions + 1) with the first matching the lattice coordinates and "
        "the last containing E0 and E1.")
    if unary_energy.ndim == 2:
        if unary_energy.shape[0] != n_nodes:
            raise ValueError(unary_energy_error + (" (# nodes = %d; # unary_array_nodes = %d)"
                                                   % (n_nodes, unary_energy.shape[0])))
        if unary_energy.shape[1] != 2:
            raise ValueError(unary_energy_error + (" (Second dimension has shape %d; 2 required.)"
                                                   % unary_energy.shape[1]))
        unary_a = unary_energy
    elif unary_energy.ndim == len(dimensions) + 1:
        for i in range(len(dimensions)):
            if unary_energy.shape[i] != dimensions[i]:
                raise ValueError(unary_energy_error + (" (dimension %d = %d; unary_array shape is %d)"
                                                       % (i, dimensions[i], unary_energy.shape[i])))
        if unary_energy.shape[-1] != 2:
            raise ValueError(unary_energy_error + (" (Last dimension has shape %d; 2 required.)"
                                                   % unary_energy.shape[-1]))
        unary_a = unary_energy.reshape(-1, 2)
    else:
        raise ValueError(unary_energy_error + " (wrong number of dimensions)")

    pairwise_energy_error = (
        "Pairwise energy must be either: (1) a 3d array with first dimension "
        "matching the number of nodes, second dimension containing the edge direction output, "
        "and third dimension containing E00, E01, E10, and E11 (with E01 being energy of "
        "source node off and dest node on.), or (2) a (#dimensions "
        "+ 2) array with last two dimensions being the same and first dimensions corresponding "
        "to the lattice coordinates. ")
    bad_edge_error = (pairwise_energy_error +
                      " (wrong number of edges specified in second-to-last dimension; got %d, needed %d).")
    bad_last_error = pairwise_energy_error + " (wrong number of pairwise energy terms; needed 4, got %d."
    ne = n_filling_edges
    if pairwise_energy.ndim == 3:
        if pairwise_energy.shape[0] != n_nodes:
            raise ValueError(pairwise_energy_error + (" (# nodes = %d; # nodes in pairwise_energy = %d)"
                                                      % (n_nodes, pairwise_energy.shape[0])))
        if pairwise_energy.shape[1] != ne:
            raise ValueError(bad_edge_error % (pairwise_energy.shape[1], ne))
        if pairwise_energy.shape[2] != 4:
            raise ValueError(bad_last_error % pairwise_energy.shape[2])
        pairwise_a = pairwise_energy
    elif pairwise_energy.ndim == len(dimensions) + 2:
        for i in range(len(dimensions)):
            if pairwise_energy.shape[i] != dimensions[i]:
                raise ValueError(pairwise_energy_error + (" (dimension %d shape is %d; pairwise_array shape is %d)"
                                                          % (i, dimensions[i], pairwise_energy.shape[i])))
        if pairwise_energy.shape[-2] != ne:
            raise ValueError(bad_edge_error % (pairwise_energy.shape[-2], ne))
        if pairwise_energy.shape[-1] != 4:
            raise ValueError(bad_last_error % pairwise_energy.shape[-1])
        pairwise_a = pairwise_energy.reshape(-1, ne, 4)
    else:
        raise ValueError(pairwise_energy_error + " (Array has wrong number of dimensions.)")
    return (np.ascontiguousarray(unary_a, dtype=np.int64),
            np.ascontiguousarray(pairwise_a, dtype=np.int64))


class LatticeFlow(object):
    """An interface to the lattice flow solvers (pyx:82-302)."""

    def __init__(self, kernel, solver, _dimensions, numerical_precision=1e-6, cap_bytes=8):
        if not lib().lf_is_valid_solver(solver.encode()):
            raise ValueError("'%s' not a valid solver name!  Valid solvers are: %s."
                             % (solver, ", ".join(validSolverNames())))
        dim = getKernelDimension(kernel)
        self._dimensions = tuple(int(d) for d in _dimensions)
        self.n_nodes = int(np.prod(self._dimensions))
        if len(self._dimensions) != dim:
            raise ValueError("Kernel '%s' is %d-dimensional; a shape of %d dimensions was provided."
                             % (kernel, dim, len(self._dimensions)))
        self.numerical_precision_factor = 1.0 / numerical_precision
        shape = (ctypes.c_int64 * dim)(*self._dimensions)
        h = ctypes.c_void_p()
        check(lib().lf_create(ctypes.byref(h), solver.encode(), kernel.encode(), dim, shape, int(cap_bytes)))
        self._h = h
        self._kernel = kernel
        self._solver = solver
        n_edges = lib().lf_n_filling_edges(h)
        e = np.empty((n_edges, dim), dtype=np.int32)
        check(lib().lf_edge_deltas(h, e.ctypes.data))
        self._edges = e

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().lf_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def ndim(self):
        return len(self._dimensions)

    @property
    def shape(self):
        return self._dimensions

    @property
    def size(self):
        return self.n_nodes

    @property
    def nFillingEdges(self):
        return int(self._edges.shape[0])

    @property
    def edges(self):
        return self._edges

    def directions(self):
        """All k directions in lattice order (edges, then their reverses)."""
        k = 2 * self.nFillingEdges
        d = np.empty((k, self.ndim), dtype=np.int32)
        check(lib().lf_directions(self._h, d.ctypes.data))
        return d

    def addEnergy(self, unary_energy, pairwise_energy):
        """pyx:157-267: scale by 1/numerical_precision, round to int64, validate, add."""
        nf = self.numerical_precision_factor
        unary_energy = np.ascontiguousarray(np.round(nf * np.asarray(unary_energy)), dtype=np.int64)
        pairwise_energy = np.ascontiguousarray(np.round(nf * np.asarray(pairwise_energy)), dtype=np.int64)
        self._add_int_energy(unary_energy, pairwise_energy)

    def addEnergyInt(self, unary_energy, pairwise_energy):
        """Add already-integral energies (no numerical_precision scaling)."""
        self._add_int_energy(np.ascontiguousarray(unary_energy, dtype=np.int64),
                             np.ascontiguousarray(pairwise_energy, dtype=np.int64))

    def _add_int_energy(self, unary_energy, pairwise_energy):
        unary_a, pairwise_a = _energy_arrays(self._dimensions, self.n_nodes, self.nFillingEdges,
                                             unary_energy, pairwise_energy)
        check(lib().lf_add_energy_potentials(self._h, unary_a.ctypes.data, pairwise_a.ctypes.data))

    def addEnergyDevice(self, unary_ptr, pairwise_ptr, stream=None):
        """Add energies already resident in device memory (int64, reference layout).

        ``stream``: the HIP stream that wrote them (default: torch's current
        stream); the library waits for it before reading."""
        if stream is None:
            stream = _torch_stream()
        check(lib().lf_add_energy_potentials_device(self._h, ctypes.c_void_p(unary_ptr),
                                                    ctypes.c_void_p(pairwise_ptr),
                                                    ctypes.c_void_p(stream)))

    def resetEnergy(self):
        check(lib().lf_reset_energies(self._h))

    def setTuning(self, rounds_per_relabel=0, local_iters=0, max_cycles=0, tile_shape=0, relabel_launches=0):
        t = _lib.lf_tuning(int(rounds_per_relabel), int(local_iters), int(max_cycles), int(tile_shape),
                           int(relabel_launches))
        check(lib().lf_set_tuning(self._h, ctypes.byref(t)))

    def getTuning(self):
        t = _lib.lf_tuning()
        check(lib().lf_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def setDebugChecks(self, on=True):
        """Check the working state's invariants after every push chunk and exact
        relabel of each solve (lf_set_debug_checks); a violation raises
        LatticeFlowError.  For small lattices: each check copies the state."""
        check(lib().lf_set_debug_checks(self._h, 1 if on else 0))

    def autotune(self, reps=1):
        """Time every tile shape / relabel schedule on the current energies and keep the fastest."""
        t = _lib.lf_tuning()
        check(lib().lf_autotune(self._h, int(reps), ctypes.byref(t)))
        return t.as_dict()

    def run(self):
        """Runs the specified solver (pyx:269-274)."""
        check(lib().lf_run(self._h))

    def runTime(self):
        """Wall seconds of the last run (pyx:277-282; the reference reports CPU time)."""
        t = ctypes.c_double()
        check(lib().lf_run_time(self._h, ctypes.byref(t)))
        return t.value

    def stats(self):
        s = _lib.lf_stats()
        check(lib().lf_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def getCut(self, which="smin"):
        """Cut as an int32 array of ``shape`` (pyx:285-302); 1 = source ("on") side.

        ``which``: "smin" (minimal source side, the default) or "smax".
        """
        sel = {"smin": _lib.LF_CUT_SMIN, "smax": _lib.LF_CUT_SMAX}[which]
        out = np.empty(self.n_nodes, dtype=np.int32)
        check(lib().lf_get_cut(self._h, out.ctypes.data, sel))
        return out.reshape(*self._dimensions)

    def getCutDevice(self, out_ptr, which="smin", stream=None):
        """Write the cut (int32, C order) into device memory; work queued on
        ``stream`` (default: torch's current stream) afterwards sees it."""
        sel = {"smin": _lib.LF_CUT_SMIN, "smax": _lib.LF_CUT_SMAX}[which]
        if stream is None:
            stream = _torch_stream()
        check(lib().lf_get_cut_device(self._h, ctypes.c_void_p(out_ptr), sel, ctypes.c_void_p(stream)))

    def flow(self):
        """Raw (doubled-unit) max-flow value of the last run."""
        f = ctypes.c_int64()
        check(lib().lf_get_flow(self._h, ctypes.byref(f)))
        return f.value

    def levels(self):
        """LatticeLevelReductions: final value r of every node (raw units), array of ``shape``."""
        out = np.empty(self.n_nodes, dtype=np.int64)
        check(lib().lf_get_levels(self._h, out.ctypes.data))
        return out.reshape(*self._dimensions)


class LatticeFlowBatch(object):
    """Many independent 2-D lattices of one kernel, solved by one device min-cut.

    The reference solves each lattice with its own ``LatticeFlow`` object
    (pyx:82-302); a batch gives the same per-lattice results -- ``flows()[i]``
    and ``getCuts()[i]`` equal ``flow()`` and ``getCut()`` of lattice i solved
    alone.  Energies are added per lattice in that lattice's own layout
    (``edges(i)`` is its edge order).
    """

    def __init__(self, kernel, shapes, numerical_precision=1e-6, cap_bytes=8):
        dim = getKernelDimension(kernel)
        if dim != 2:
            raise ValueError("Kernel '%s' is %d-dimensional; batches take 2-D kernels." % (kernel, dim))
        self._shapes = [tuple(int(d) for d in s) for s in shapes]
        for s in self._shapes:
            if len(s) != 2:
                raise ValueError("every lattice of a batch is 2-D; got shape %s" % (s,))
        arr = np.ascontiguousarray(np.array(self._shapes, dtype=np.int64).reshape(-1))
        h = ctypes.c_void_p()
        check(lib().lf_batch_create(ctypes.byref(h), kernel.encode(), len(self._shapes), arr.ctypes.data,
                                    int(cap_bytes)))
        self._h = h
        self._kernel = kernel
        self.numerical_precision_factor = 1.0 / numerical_precision
        self._kp = lib().lf_n_filling_edges(h)
        self._offsets = np.concatenate([[0], np.cumsum([a * b for a, b in self._shapes])]).astype(np.int64)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().lf_destroy(h)
            except Exception:
                pass
            self._h = None

    def __len__(self):
        return len(self._shapes)

    @property
    def shapes(self):
        return list(self._shapes)

    @property
    def size(self):
        return int(self._offsets[-1])

    def canvas(self):
        rc = np.empty(2, np.int64)
        check(lib().lf_batch_canvas(self._h, rc.ctypes.data))
        return int(rc[0]), int(rc[1])

    def edges(self, i):
        e = np.empty((self._kp, 2), dtype=np.int32)
        check(lib().lf_batch_edge_deltas(self._h, int(i), e.ctypes.data))
        return e

    def addEnergy(self, i, unary_energy, pairwise_energy):
        nf = self.numerical_precision_factor
        self.addEnergyInt(i, np.round(nf * np.asarray(unary_energy)), np.round(nf * np.asarray(pairwise_energy)))

    def addEnergyInt(self, i, unary_energy, pairwise_energy):
        i = int(i)
        if not 0 <= i < len(self._shapes):
            raise IndexError("batch lattice index %d out of range" % i)
        n = self._shapes[i][0] * self._shapes[i][1]
        u = np.ascontiguousarray(unary_energy, dtype=np.int64)
        p = np.ascontiguousarray(pairwise_energy, dtype=np.int64)
        if u.size != 2 * n:
            raise ValueError("unary energy of lattice %d must hold %d x 2 values" % (i, n))
        if p.size != n * self._kp * 4:
            raise ValueError("pairwise energy of lattice %d must hold %d x %d x 4 values" % (i, n, self._kp))
        check(lib().lf_batch_add_energy_potentials(self._h, i, u.ctypes.data, p.ctypes.data))

    def addEnergyDevice(self, i, unary_ptr, pairwise_ptr, stream=None):
        """Energies of lattice i from device memory; ``stream`` (default: torch's
        current stream) is the stream that wrote them."""
        if stream is None:
            stream = _torch_stream()
        check(lib().lf_batch_add_energy_potentials_device(self._h, int(i), ctypes.c_void_p(unary_ptr),
                                                          ctypes.c_void_p(pairwise_ptr),
                                                          ctypes.c_void_p(stream)))

    def resetEnergy(self):
        check(lib().lf_reset_energies(self._h))

    def setTuning(self, rounds_per_relabel=0, local_iters=0, max_cycles=0, tile_shape=0, relabel_launches=0):
        t = _lib.lf_tuning(int(rounds_per_relabel), int(local_iters), int(max_cycles), int(tile_shape),
                           int(relabel_launches))
        check(lib().lf_set_tuning(self._h, ctypes.byref(t)))

    def getTuning(self):
        t = _lib.lf_tuning()
        check(lib().lf_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def setDebugChecks(self, on=True):
        """Check the working state's invariants after every push chunk and exact
        relabel of each solve (lf_set_debug_checks); a violation raises
        LatticeFlowError.  For small lattices: each check copies the state."""
        check(lib().lf_set_debug_checks(self._h, 1 if on else 0))

    def autotune(self, reps=1):
        t = _lib.lf_tuning()
        check(lib().lf_autotune(self._h, int(reps), ctypes.byref(t)))
        return t.as_dict()

    def run(self):
        check(lib().lf_run(self._h))

    def runTime(self):
        t = ctypes.c_double()
        check(lib().lf_run_time(self._h, ctypes.byref(t)))
        return t.value

    def stats(self):
        s = _lib.lf_stats()
        check(lib().lf_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def flows(self):
        """Raw (doubled-unit) max-flow value of every lattice of the last run."""
        out = np.empty(len(self._shapes), dtype=np.int64)
        check(lib().lf_batch_get_flows(self._h, out.ctypes.data))
        return out

    def getCutsFlat(self, which="smin"):
        sel = {"smin": _lib.LF_CUT_SMIN, "smax": _lib.LF_CUT_SMAX}[which]
        out = np.empty(self.size, dtype=np.int32)
        check(lib().lf_get_cut(self._h, out.ctypes.data, sel))
        return out

    def getCuts(self, which="smin"):
        """Cut of every lattice, a list of int32 arrays of each lattice's shape."""
        flat = self.getCutsFlat(which)
        return [flat[self._offsets[i]:self._offsets[i + 1]].reshape(s) for i, s in enumerate(self._shapes)]


def _benchmark_energies(X, kernel, edges):
    """Energies of the reference's image benchmark (pyx:305-398), vectorised.

    2-D kernels: one node per pixel with the pixel's RGB colour; 3-D kernels:
    nodes (x, y, z) over an H x W x W volume with colour (X[x, y] + X[x, z]) / 2.
    E0 = 50000 on pure black nodes, E1 = 50000 on pure white ones; every
    in-lattice edge e gets E01 = E10 = exp(-|c1 - c2|^2) / |delta_e|^2, and the
    smallest E01 of the whole array (0 wherever an edge leaves the lattice) is
    subtracted from the E01 entries (pyx:393-394)."""
    X = np.asarray(X, dtype=np.float64)[:, :, :3]
    dim = getKernelDimension(kernel)
    H, W = X.shape[0], X.shape[1]
    if dim == 2:
        shape = (H, W)
        col = X
    elif dim == 3:
        shape = (H, W, W)
        col = (X[:, :, None, :] + X[:, None, :, :]) / 2
    else:
        raise ValueError("Benchmark for dimension %d not implemented (kernel=%s)" % (dim, kernel))
    n = int(np.prod(shape))
    flat = col.reshape(n, 3)
    unary = np.zeros((n, 2))
    black = (flat[:, 0] == flat[:, 1]) & (flat[:, 1] == flat[:, 2]) & (flat[:, 2] == 0)
    white = (flat[:, 0] == flat[:, 1]) & (flat[:, 1] == flat[:, 2]) & (flat[:, 2] == 1)
    unary[black, 0] = 50000
    unary[white, 1] = 50000
    binary = np.zeros((n, len(edges), 4))
    for ei, dlt in enumerate(np.asarray(edges)):
        scale = 1.0 / float(np.sum(dlt.astype(np.float64) ** 2))
        src = tuple(slice(max(0, -int(d)), shape[a] - max(0, int(d))) for a, d in enumerate(dlt))
        dst = tuple(slice(max(0, int(d)), shape[a] + min(0, int(d))) for a, d in enumerate(dlt))
        c1 = col[src]
        c2 = col[dst]
        d = np.exp(-np.sum((c1 - c2) ** 2, axis=-1)) * scale
        plane = np.zeros(shape)
        plane[src] = d
        binary[:, ei, 1] = plane.reshape(-1)
        binary[:, ei, 2] = plane.reshape(-1)
    binary[:, :, 1:2] -= binary[:, :, 1].min()
    return shape, unary, binary


def _runBenchmark(image_file, kernel, solver, check_result=False):
    """pyx:305-435: segment an RGB image (a PNG path, read with matplotlib's
    imread as the reference does, or an array H x W x 3+) with ``solver``, print
    the result and return runTime().  With ``check_result`` the same energies
    are solved by the CPU solver "BKGCEnergyMinimizer" and the cuts compared.
    The printed solution hash is md5 of the cut's bytes (the reference hashes a
    Python 2 pickle of it)."""
    import hashlib
    if isinstance(image_file, str):
        from matplotlib.pyplot import imread
        Xo = imread(image_file)
        if not Xo.size:
            raise IOError("Error loading image %s." % image_file)
    else:
        Xo = np.asarray(image_file)
    if Xo.ndim != 3:
        raise IOError("Error loading image %s; need RGB images." % (image_file if isinstance(image_file, str) else "array"))
    lf = LatticeFlow(kernel, solver, (Xo.shape[0], Xo.shape[1]) if getKernelDimension(kernel) == 2
                     else (Xo.shape[0], Xo.shape[1], Xo.shape[1]))
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
        lf2 = LatticeFlow(kernel, "BKGCEnergyMinimizer", shape)
        lf2.addEnergy(unary, binary)
        lf2.run()
        print("BKGC Solver ran in %f seconds." % lf2.runTime())
        cut2 = lf2.getCut().ravel()
        if (cut != cut2).sum() != 0:
            print("Solution different: %d falsely 0, %d falsely 1"
                  % (((cut == 0) & (cut2 == 1)).sum(), ((cut == 1) & (cut2 == 0)).sum()))
        else:
            print("Solutions exactly match.")
    return runtime


def _tv_single(X, lam, kernel):
    X = np.ascontiguousarray(X, dtype=np.float64)
    if X.ndim != 2:
        raise ValueError("calculate2dTV needs a 2-d array")
    nx, ny = X.shape
    R = np.empty((nx, ny), dtype=np.float64)
    st = _lib.lf_stats()
    check(lib().lf_tv_single(kernel.encode(), nx, ny, X.ctypes.data, float(lam), R.ctypes.data,
                             ctypes.byref(st)))
    return R, st.as_dict()


def calculate2dTV(X, lam, kernel="Star2d_16", return_stats=False):
    """pyx:437-454: exact 2-D TV of X at one lambda, times lambda (as the reference's
    Python layer returns it).  ``kernel`` (new, keyword) selects the geocut
    kernel; the reference hard-wires Star2d_16."""
    R, st = _tv_single(X, lam, kernel)
    R = R * lam
    return (R, st) if return_stats else R


def _tv_path(X, lambdas, kernel):
    X = np.ascontiguousarray(X, dtype=np.float64)
    if X.ndim != 2:
        raise ValueError("calculate2dTVPath needs a 2-d array")
    lam = np.ascontiguousarray(lambdas, dtype=np.float64).reshape(-1)
    nx, ny = X.shape
    R = np.empty((lam.size, nx, ny), dtype=np.float64)
    st = _lib.lf_stats()
    check(lib().lf_tv_path(kernel.encode(), nx, ny, X.ctypes.data, lam.size, lam.ctypes.data,
                           R.ctypes.data, ctypes.byref(st)))
    return R, st.as_dict()


def _tv_path_guided(X, lambdas, kernel, pred, delta, max_repairs=8):
    """Exact TV path from an initial partition by the device predictions ``pred``
    (a float32 CUDA tensor of shape (len(lambdas),) + X.shape).  Same output as
    ``_tv_path``; returns (R, stats, guided_stats)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    lam = np.ascontiguousarray(lambdas, dtype=np.float64).reshape(-1)
    nx, ny = X.shape
    if tuple(pred.shape) != (lam.size, nx, ny) or str(pred.dtype) != "torch.float32" or not pred.is_cuda:
        raise ValueError("pred must be a float32 device tensor of shape %s" % ((lam.size, nx, ny),))
    pred = pred.contiguous()
    R = np.empty((lam.size, nx, ny), np.float64)
    st = _lib.lf_stats()
    gs = _lib.lf_guided_stats()
    check(lib().lf_tv_path_guided(kernel.encode(), nx, ny, X.ctypes.data, lam.size, lam.ctypes.data,
                                  ctypes.c_void_p(pred.data_ptr()), float(delta), int(max_repairs),
                                  R.ctypes.data, ctypes.byref(st), ctypes.byref(gs),
                                  ctypes.c_void_p(_torch_stream(pred.device))))
    return R, st.as_dict(), gs.as_dict()


def calculate2dTVPath(X, lambdas, kernel="Star2d_16", return_stats=False):
    """pyx:457-484: the exact regularization path of X at ``lambdas``.  As in the
    reference's Python layer, every slice is divided by the last lambda of the
    list (its loop index is stale there)."""
    R, st = _tv_path(X, lambdas, kernel)
    lam = np.asarray(lambdas, dtype=np.float64).reshape(-1)
    if lam.size:
        R = R / lam[-1]
    return (R, st) if return_stats else R


class TVRegPath(object):
    """The event-driven regularization path of the 2-d image X over lambda in
    [0, lambda_max] (TVSolver::run, tv_solver.hpp:114-149,300-642): the exact
    solution at lambda_max, then closed-form joins and device-verified splits
    of its level sets down to lambda = 0.  Built once; ``path(lambdas)`` gives
    the slices at any lambdas in that range (getRegularizationPath,
    tv_solver.hpp:181-244), the same values as ``calculate2dTVPath``'s C++
    counterpart ``_tv_path`` with lambda_max the largest lambda of the list.
    ``host_max``: level sets of at most this many nodes are checked on the
    host, larger ones on the device (0: the library's default; negative: all
    on the device)."""

    def __init__(self, X, lambda_max, kernel="Star2d_16", host_max=0):
        X = np.ascontiguousarray(X, dtype=np.float64)
        if X.ndim != 2:
            raise ValueError("TVRegPath needs a 2-d array")
        self._shape = X.shape
        self._p = None
        h = ctypes.c_void_p()
        check(lib().lf_tvpath_create(ctypes.byref(h), kernel.encode(), X.shape[0], X.shape[1], X.ctypes.data,
                                     float(lambda_max), int(host_max)))
        self._p = h

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value:
            try:
                lib().lf_tvpath_destroy(p)
            except Exception:
                pass
            self._p = None

    def __call__(self, lambdas):
        lam = np.ascontiguousarray(lambdas, dtype=np.float64).reshape(-1)
        R = np.empty((lam.size,) + self._shape, dtype=np.float64)
        check(lib().lf_tvpath_eval(self._p, lam.size, lam.ctypes.data, R.ctypes.data))
        return R

    def stats(self):
        st = _lib.lf_tvpath_stats()
        check(lib().lf_tvpath_get_stats(self._p, ctypes.byref(st)))
        return st.as_dict()


def calculate2dTVPathEvents(X, lambdas, kernel="Star2d_16", return_stats=False):
    """calculate2dTV(nx, ny, f, vector<double> lambda) (tv_solver.hpp:659-681)
    through the event-driven path: one path up to the largest lambda, read at
    every lambda of the list.  Slices as ``_tv_path`` gives them (before the
    Python layer's division by the last lambda)."""
    lam = np.ascontiguousarray(lambdas, dtype=np.float64).reshape(-1)
    X = np.ascontiguousarray(X, dtype=np.float64)
    if lam.size == 0:
        R = np.empty((0,) + X.shape, np.float64)
        return (R, {}) if return_stats else R
    path = TVRegPath(X, max(float(np.maximum(lam, 0).max()), 0.0), kernel)
    R = path(np.maximum(lam, 0))
    return (R, path.stats()) if return_stats else R


def verify_state(dhw, dirs, red0, alpha0, res, ex, mbox=None, mask=None, h=None, labels_exact=False):
    """The library's state checks (lf_verify_state) on host arrays in the device
    engine's layout: dhw = (D, H, W); dirs (k, 3) int offsets in lattice order;
    red0 (N,) and alpha0 (k, N) int64 base network; res (k, N) and ex (N,) of
    int32 or int64; mbox (P, k, N) flow in flight or None; mask (N,) unsigned
    arc bits or None; h (N,) int32 labels or None.  Returns None when every
    invariant holds, else the library's description of the first violation."""
    import numpy as np
    dhw = np.ascontiguousarray(dhw, np.int64)
    dirs = np.ascontiguousarray(dirs, np.int32).reshape(-1, 3)
    k = dirs.shape[0]
    red0 = np.ascontiguousarray(red0, np.int64)
    alpha0 = np.ascontiguousarray(alpha0, np.int64)
    res = np.ascontiguousarray(res)
    cap_bytes = res.dtype.itemsize
    ex = np.ascontiguousarray(ex, res.dtype)
    planes = 0
    if mbox is not None:
        mbox = np.ascontiguousarray(mbox, res.dtype)
        planes = mbox.shape[0]
    if mask is not None:
        mask = np.ascontiguousarray(mask)
    if h is not None:
        h = np.ascontiguousarray(h, np.int32)
    ptr = lambda a: None if a is None else a.ctypes.data  # noqa: E731
    rc = lib().lf_verify_state(ptr(dhw), k, ptr(dirs), ptr(red0), ptr(alpha0), cap_bytes, ptr(res), ptr(ex),
                               ptr(mbox), planes, ptr(mask), 0 if mask is None else mask.dtype.itemsize, ptr(h),
                               1 if labels_exact else 0)
    if rc == _lib.LF_OK:
        return None
    if rc == _lib.LF_ERR_STATE:
        return lib().lf_last_error().decode()
    check(rc)

