"""tom_amd — MI355X (gfx950) implementation of the TOM Toolbox template
matching / subtomogram scoring hot path, behind a C ABI (include/tomhip.h)
and a MATLAB-shaped Python interface (tom_amd.api)."""
from ._lib import LIB_PATH, TomHipError, load  # noqa: F401

__all__ = ["LIB_PATH", "TomHipError", "load"]
