"""MI355X-native lattice push-relabel min-cut (hoytak/latticeflow's hot path).

The product is the HIP library ``liblatticeflow_amd.so`` behind the C ABI in
``include/latticeflow_amd.h``; this package holds its Python host mirror
(``pylatticeflow``), the synthetic workload generators and the build script.
"""
from .pylatticeflow import (LatticeFlow, LatticeFlowBatch, TVRegPath, calculate2dTV, calculate2dTVPath,  # noqa: F401
                            calculate2dTVPathEvents, getKernelDimension, validKernelNames, validSolverNames)

__all__ = ["LatticeFlow", "LatticeFlowBatch", "TVRegPath", "calculate2dTV", "calculate2dTVPath",
           "calculate2dTVPathEvents", "getKernelDimension", "validKernelNames", "validSolverNames"]
