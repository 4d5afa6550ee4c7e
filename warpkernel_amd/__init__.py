"""warpkernel_amd -- WarpKernel FE SpMV (WPK1/WPK2) for MI355X (gfx950).

The product is libwarpkernel.so (C ABI: include/warpkernel_c.h).  This package
binds it with ctypes:

  warpkernel_amd.structure / structure2 / engine   the reference's class API
  warpkernel_amd.fe                                FE workload generators
  warpkernel_amd.autograd                          torch.autograd.Function
"""
from ._capi import WKError, default_launch, lib  # noqa: F401
from .api import batch, engine, graph, structure, structure2  # noqa: F401

__all__ = ["structure", "structure2", "engine", "batch", "graph", "WKError", "default_launch", "lib"]
