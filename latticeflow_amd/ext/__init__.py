"""The Cython extension ``pylatticeflow`` (a Python 3 re-creation of the
reference's extensions/python/pylatticeflow.pyx over the C++ interface headers
of include/latticeflow_amd/).  Built in place by ``latticeflow_amd.build``;
a code base that does ``import pylatticeflow`` puts this directory on its path.

PyTorch ships its own HIP runtime under the same soname as ROCm's.  Whichever
is loaded first serves the whole process; two copies in one process do not
see the device.  PyTorch is therefore loaded and initialised before the
extension (and with it the library) is imported, as the ctypes layer does.  A
program that imports ``pylatticeflow`` directly and uses PyTorch as well must
import and initialise PyTorch first."""
try:
    import torch as _torch

    _torch.cuda.is_available()
except ImportError:
    pass
