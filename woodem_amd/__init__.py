"""woodem_amd: the Woo DEM timestep (InsertionSortCollider, ContactLoop with L6Geom
functors and FrictPhys/Hertz laws, Leapfrog, EnergyTracker) on AMD MI355X.

The compute path is libwoodem_hip.so (hand-written HIP kernels for gfx950)
behind a C ABI (include/woodem_hip.h).  This package mirrors the reference's
Python scripting surface (woo.core, woo.dem, woo.utils, woo.pack's generators) for that path.
"""
from . import _native  # noqa: F401  (fails loudly if the library is missing, on first use)
from .core import Cell, EnergyTracker, Scene  # noqa: F401
from .dem import *  # noqa: F401,F403
from . import utils  # noqa: F401
from . import soa  # noqa: F401
from . import pack  # noqa: F401
