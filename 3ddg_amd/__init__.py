"""MI355X (gfx950) implementation of drewalan/3DDG's per-RK-stage right-hand side and update.

The package directory name starts with a digit, so import it with
``importlib.import_module("3ddg_amd")``. Submodules:
  dg3d     ctypes binding of include/dg3d.h (libdg3d.so)
  problem  host-side problem setup (mesh, material, sponge, IC, halos, source, workloads)
  grid     the reference's Grid/Block method names over the C ABI (Grid.h, Block.h)
The reference's parameter-file format (driver.cpp:129-196) is read by the C++ driver
(csrc/host/params.cpp), not from Python.
"""
from . import dg3d  # noqa: F401
from . import problem  # noqa: F401
from . import grid  # noqa: F401
