"""MI355X-native implementation of DIABLO 3's channel RK3/Crank-Nicolson substep.

The compute path is the HIP library ``libdgpu.so`` (sources in ``csrc/``,
C-ABI in ``include/diablo_gpu.h``); ``binding`` drives it through ctypes with
the reference's procedure names (rk_chan_1, rem_div_chan, ...).
"""
from .binding import ChannelSolver, DgpuError, Params, load_library, FIELD_IDS  # noqa: F401
from .grid import create_grid, metrics  # noqa: F401
from . import workloads  # noqa: F401
