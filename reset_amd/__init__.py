"""RESET-MI355X: hash-based set intersection on gfx950 (HIP) behind a C ABI.

``reset_amd.lib`` binds the C ABI (include/reset_amd.h); ``reset_amd.api``
is the torch-tensor host API.  Importing the package does not touch the GPU.
"""
from .lib import (BUILD_BOTH, BUILD_HIERARCHICAL, BUILD_NORMAL, HIERARCHICAL, NORMAL, ResetError,  # noqa: F401
                  load)

__all__ = ["load", "ResetError", "NORMAL", "HIERARCHICAL", "BUILD_NORMAL", "BUILD_HIERARCHICAL", "BUILD_BOTH"]
