"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and its name / kernel / edge-order tables agree with the
oracle's restatement of the reference generator (install/code_setup.py,
kernellattice.hpp:61-96).  No compute call is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    syms = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            syms |= set(re.findall(r"\b(lf_[a-z0-9_]+)\s*\(", txt))
    return syms


def test_library_exports_all_declared_symbols():
    from latticeflow_amd import _lib
    L = _lib.lib()
    declared = _header_symbols()
    assert declared, "no declarations found"
    for s in declared:
        assert hasattr(L, s), s
    # the Python binding covers the header exactly
    assert declared == set(_lib.declared_symbols())


def test_names_and_dimensions(oracle):
    import latticeflow_amd as L
    fam = oracle.kernel_family()
    names = set(L.validKernelNames())
    assert names == set(fam) | {"Full3d_26"}
    for n, (nd, _) in fam.items():
        assert L.getKernelDimension(n) == nd
    assert L.getKernelDimension("Full3d_26") == 3
    assert set(L.validSolverNames()) == {"LatticeEnergyMinimizer", "LatticeLevelReductions", "BKGCEnergyMinimizer"}
    with pytest.raises(ValueError):
        L.getKernelDimension("Full2d_5")
    with pytest.raises(ValueError):
        L.LatticeFlow("Full2d_4", "NoSuchSolver", (4, 4))
    with pytest.raises(ValueError):
        L.LatticeFlow("Full2d_4", "LatticeEnergyMinimizer", (4, 4, 4))


def test_kernel_offsets_and_geocut_match_oracle(oracle):
    from latticeflow_amd._lib import lib
    for n, (nd, edges) in oracle.kernel_family().items():
        k = lib().lf_kernel_size(n.encode())
        assert k == len(edges)
        offs = np.empty((k, nd), np.int32)
        w = np.empty(k)
        assert lib().lf_kernel_offsets(n.encode(), offs.ctypes.data, w.ctypes.data) == 0
        assert [tuple(r) for r in offs.tolist()] == list(edges), n
        _, wo = oracle.geocut_weights(edges)
        assert np.array_equal(w, np.array(wo)), n


@pytest.mark.parametrize("kernel,shape", [("Full2d_8", (7, 9)), ("Full2d_4", (1, 5)), ("Star2d_16", (6, 11)),
                                          ("Star3d_26", (3, 4, 5)), ("Full2d_48", (20, 20)), ("Full1d_4", (9,))])
def test_edge_order_matches_oracle(oracle, kernel, shape):
    from latticeflow_amd import workloads as W
    a, wa = W.edge_order(kernel, shape)
    b, wb = oracle.edge_order(kernel, shape)
    assert np.array_equal(a, b) and np.array_equal(wa, wb)


def test_last_error_is_thread_local_string():
    from latticeflow_amd._lib import lib
    assert lib().lf_kernel_dimension(b"nope") == -1
    assert b"nope" in lib().lf_last_error()
    assert lib().lf_is_valid_kernel(None) == 0


def test_batch_create_rejects_bad_arguments():
    """lf_batch_create validates its arguments before touching the device."""
    from latticeflow_amd._lib import lib, LF_ERR_BAD_NAME, LF_ERR_UNSUPPORTED, LF_ERR_BAD_ARG
    h = ctypes.c_void_p()
    shapes = (ctypes.c_int64 * 4)(10, 20, 30, 40)
    assert lib().lf_batch_create(ctypes.byref(h), b"Nope2d_8", 2, shapes, 4) == LF_ERR_BAD_NAME
    assert lib().lf_batch_create(ctypes.byref(h), b"Full3d_6", 2, shapes, 4) == LF_ERR_UNSUPPORTED
    assert b"2-D" in lib().lf_last_error()
    assert lib().lf_batch_create(ctypes.byref(h), b"Full2d_8", 0, shapes, 4) == LF_ERR_BAD_ARG
    assert lib().lf_batch_create(ctypes.byref(h), b"Full2d_8", 2, shapes, 3) == LF_ERR_BAD_ARG
    bad = (ctypes.c_int64 * 4)(10, 0, 30, 40)
    assert lib().lf_batch_create(ctypes.byref(h), b"Full2d_8", 2, bad, 4) == LF_ERR_BAD_ARG
    assert not h.value
