"""The Cython extension ``pylatticeflow`` (a Python 3 re-creation of the
reference's extensions/python/pylatticeflow.pyx over the C++ interface headers
of include/latticeflow_amd/).  Built in place by ``latticeflow_amd.build``;
a code base that does ``import pylatticeflow`` puts this directory on its path."""
