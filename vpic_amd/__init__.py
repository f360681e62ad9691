"""vpic_amd -- MI355X (gfx950) implementation of VPIC's per-step
particle-in-cell hot path behind a C-ABI (include/vpic_hip.h)."""
from . import types  # noqa: F401
