"""bksp_amd -- MI355X-native (gfx950) drop-in for the Krylov hot path of bksp.

The product is libbksp_amd.so: hand-written HIP kernels behind a C-ABI, and C
host code exporting the reference's bksp.h entry points. This package only
loads it (``_lib``), mirrors bksp.h in Python (``bksp``), exposes the
device-resident extension API (``device``) and the seeded input generators
(``gen``).
"""
from . import bksp, device, gen  # noqa: F401
from ._lib import LIB_PATH, load  # noqa: F401
