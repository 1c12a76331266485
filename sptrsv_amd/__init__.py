"""MI355X (gfx950) sparse lower-triangular solver: HIP kernels behind a C-ABI
(include/sptrsv_amd.h), with the host interface of HCL-Fing/SPTRSV mirrored
in ``sptrsv_amd.interface``."""
from ._lib import SptrsvError, SptrsvUnavailable, lib  # noqa: F401

__all__ = ["SptrsvError", "SptrsvUnavailable", "lib", "interface"]
