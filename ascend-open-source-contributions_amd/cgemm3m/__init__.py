"""MI355X (gfx950) aclblasCgemm3m: Python side of the C ABI.

The compute path is libaclblas_cgemm3m.so (hand-written HIP kernels); this package only
binds it (ctypes) and adds the torch autograd wrapper. torch supplies device memory and
streams.
"""
from . import _lib
from .api import (Cgemm3mError, Handle, cgemm, cgemm3m, config, grouped, grouped_workspace_size, residual,
                  strided_batched, torch_ptr, workspace_size)

__all__ = ["Cgemm3mError", "Handle", "cgemm", "cgemm3m", "config", "grouped", "grouped_workspace_size", "residual",
           "strided_batched", "torch_ptr", "workspace_size", "_lib"]
