"""MI355X-native (gfx950) fractional-step TBL time step.

The compute path is libtbl_hip.so (hand-written HIP, C ABI in include/tbl.h);
``turbulentbl_amd.tbl`` binds it, ``turbulentbl_amd.inflow`` builds seeded
initial fields and recycling state, ``turbulentbl_amd.fortran`` holds the
iso_c_binding module and driver.
"""
from . import inflow  # noqa: F401

__all__ = ["inflow", "tbl"]
