"""Python bindings (ctypes) over the C ABI in include/dgc_hip.h.

The compute path is the HIP library lib/libdgc_amd.so; this module only
moves host arrays in and out and forwards calls. There is no CPU fallback:
if the library is missing, every entry point raises.
"""
from .core import (  # noqa: F401
    DGCError,
    Context,
    DeviceArray,
    Mesh,
    Discretization,
    MaxwellTM,
    Advection,
    ref_operators,
    lib,
    lib_path,
    node_to_cell_major,
    cell_to_node_major,
)
from .ipdg import IPDG, poisson_rhs  # noqa: F401
from .graddiv import GradDiv  # noqa: F401
from .stokes import Stokes  # noqa: F401
from .advection1d import Advection1D, write_plot_file  # noqa: F401
from .mpnn import LearnedPreconditioner  # noqa: F401
from .ensemble import Ensemble  # noqa: F401
