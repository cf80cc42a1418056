"""Hash-grid encoder: Python side of the ``nect_grid_*`` entry points.

``GridDesc`` holds the level table the library computes (nect_grid_init) for
tiny-cuda-nn style ``HashGrid`` settings (nect/config.py:86-112).  The
functions take GPU tensors and launch on torch's current stream.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


class GridDesc:
    """Level table of a multiresolution hash grid (see include/nect_amd.h)."""

    def __init__(self, n_dims: int, n_levels: int, n_features_per_level: int, log2_hashmap_size: int,
                 base_resolution: int, per_level_scale: float):
        self.struct = _lib.NectGridDesc()
        _lib.call("nect_grid_init", ctypes.byref(self.struct), int(n_dims), int(n_levels), int(n_features_per_level),
                  int(log2_hashmap_size), int(base_resolution), float(per_level_scale))
        self.n_dims = int(n_dims)
        self.n_levels = int(n_levels)
        self.n_features = int(n_features_per_level)
        self.n_params = int(self.struct.n_params)
        self.n_output_dims = self.n_levels * self.n_features

    def struct_ref(self):
        """ctypes pointer to the descriptor, for the C entry points."""
        return ctypes.byref(self.struct)

    @property
    def scale(self):
        return [self.struct.scale[i] for i in range(self.n_levels)]

    @property
    def resolution(self):
        return [self.struct.resolution[i] for i in range(self.n_levels)]

    @property
    def level_size(self):
        return [self.struct.level_size[i] for i in range(self.n_levels)]

    @property
    def offset(self):
        return [self.struct.offset[i] for i in range(self.n_levels + 1)]


def _rows(x: torch.Tensor, min_cols: int) -> tuple[int, int]:
    if x.dim() != 2 or x.stride(1) != 1 or x.shape[1] < min_cols:
        raise ValueError(f"expected a [n, >={min_cols}] tensor with unit column stride")
    return x.shape[0], x.stride(0)


def index(desc: GridDesc, x: torch.Tensor):
    """(index uint32 as int32 [n, L, 2^D], weight [n, L, 2^D]) of the corners of x [n, D]."""
    n, xs = _rows(x, desc.n_dims)
    K = 1 << desc.n_dims
    idx = torch.empty((n, desc.n_levels, K), dtype=torch.int32, device=x.device)
    w = torch.empty((n, desc.n_levels, K), dtype=torch.float32, device=x.device)
    _lib.call("nect_grid_index", ctypes.byref(desc.struct), x.data_ptr() if n else None, n, xs,
              _lib.ptr(idx), _lib.ptr(w), _lib.stream_of(x.device))
    return idx, w


def encode(desc: GridDesc, table: torch.Tensor, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """Features [n, L*F] in table.dtype of x [n, D] (float32, in [0, 1])."""
    if x.dtype != torch.float32:
        raise ValueError("grid inputs must be float32")
    if table.numel() != desc.n_params:
        raise ValueError(f"table has {table.numel()} elements, the grid needs {desc.n_params}")
    n, xs = _rows(x, desc.n_dims)
    if not x.is_cuda:
        raise ValueError("nect_amd kernels take GPU tensors only")
    if out is None:
        out = torch.empty((n, desc.n_output_dims), dtype=table.dtype, device=x.device)
    _lib.call("nect_grid_encode", ctypes.byref(desc.struct), _lib.dtype_code(table.dtype), _lib.ptr(table),
              x.data_ptr() if n else None, n, xs, out.data_ptr() if n else None, out.stride(0),
              _lib.stream_of(x.device))
    return out


def backward(desc: GridDesc, x: torch.Tensor, dout: torch.Tensor, grad: torch.Tensor) -> torch.Tensor:
    """grad[n_params] (float32) += d(sum(dout * encode))/d(table)."""
    if dout.dtype != torch.float32 or grad.dtype != torch.float32:
        raise ValueError("grid backward takes float32 dout and grad")
    n, xs = _rows(x, desc.n_dims)
    if not x.is_cuda:
        raise ValueError("nect_amd kernels take GPU tensors only")
    _lib.call("nect_grid_backward", ctypes.byref(desc.struct), x.data_ptr() if n else None, n, xs,
              dout.data_ptr() if n else None, dout.stride(0), _lib.ptr(grad), _lib.stream_of(x.device))
    return grad
