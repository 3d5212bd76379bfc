"""fieldgrid_amd: an MI355X-native (gfx950 HIP) Yee-grid FDTD engine behind the C-ABI of
include/fieldgrid.h, with a Python mirror of sims-field's FDTD API in `fdtd`."""
from . import abi, fdtd  # noqa: F401
