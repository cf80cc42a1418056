"""Fully fused MLP: Python side of the ``nect_mlp_*`` entry points.

Mirrors tiny-cuda-nn's ``FullyFusedMLP`` network config (nect/config.py:
125-132): bias-free, ``n_hidden_layers`` layers of 64 neurons, activation
names as the reference's config accepts them (nect/config.py:704-714).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib

ACTIVATIONS = {
    "none": 0, "relu": 1, "leakyrelu": 2, "exponential": 3, "sine": 4, "sigmoid": 5, "squareplus": 6,
    "softplus": 7, "tanh": 8,
}


def activation_code(name: str | None) -> int:
    key = "none" if name is None else str(name).casefold()
    if key not in ACTIVATIONS:
        raise ValueError(f"unsupported activation '{name}'; supported: {sorted(ACTIVATIONS)}")
    return ACTIVATIONS[key]


class MlpDesc:
    """Shape of a bias-free MLP and its flat weight layout (see include/nect_amd.h)."""

    def __init__(self, n_input: int, n_output: int, n_neurons: int = 64, n_hidden_layers: int = 2,
                 activation: str = "ReLU", output_activation: str = "None"):
        self.struct = _lib.NectMlpDesc()
        _lib.call("nect_mlp_init", ctypes.byref(self.struct), int(n_input), int(n_output), int(n_neurons),
                  int(n_hidden_layers), activation_code(activation), activation_code(output_activation))
        self.n_input = int(n_input)
        self.n_output = int(n_output)
        self.width = int(n_neurons)
        self.n_hidden_layers = int(n_hidden_layers)
        self.n_input_padded = self.struct.n_input_padded
        self.n_output_padded = self.struct.n_output_padded
        self.activation = activation
        self.output_activation = output_activation
        self.n_params = int(self.struct.n_params)

    def struct_ref(self):
        """ctypes pointer to the descriptor, for the C entry points."""
        return ctypes.byref(self.struct)

    def layer_shapes(self) -> list[tuple[int, int]]:
        """(out, in) of each weight matrix in storage order."""
        shapes = [(self.width, self.n_input_padded)]
        shapes += [(self.width, self.width)] * (self.n_hidden_layers - 1)
        shapes.append((self.n_output_padded, self.width))
        return shapes

    def split(self, flat: torch.Tensor) -> list[torch.Tensor]:
        """Views of the weight matrices inside a flat parameter vector."""
        out, o = [], 0
        for r, c in self.layer_shapes():
            out.append(flat[o:o + r * c].view(r, c))
            o += r * c
        return out


def forward(desc: MlpDesc, weights: torch.Tensor, x: torch.Tensor, out: torch.Tensor | None = None,
            n_tiles_dev: torch.Tensor | None = None) -> torch.Tensor:
    """y [n, n_output] float32 = MLP(x) with x and weights in the same dtype."""
    if weights.dtype != x.dtype:
        raise ValueError("weights and inputs must share a dtype")
    if weights.numel() != desc.n_params:
        raise ValueError(f"weights have {weights.numel()} elements, the MLP needs {desc.n_params}")
    n = x.shape[0]
    if out is None:
        out = torch.empty((n, desc.n_output), dtype=torch.float32, device=x.device)
    if n == 0:
        return out
    if not x.is_cuda or x.stride(1) != 1:
        raise ValueError("x must be a GPU tensor with unit column stride")
    _lib.call("nect_mlp_forward", ctypes.byref(desc.struct), _lib.dtype_code(x.dtype), _lib.ptr(weights),
              x.data_ptr(), n, x.stride(0), None, _lib.ptr(n_tiles_dev), out.data_ptr(), out.stride(0),
              _lib.stream_of(x.device))
    return out


def workspace_bytes(desc: MlpDesc, dtype: torch.dtype, n: int) -> int:
    return int(_lib.lib().nect_mlp_backward_workspace(ctypes.byref(desc.struct), _lib.dtype_code(dtype), int(n)))


def backward(desc: MlpDesc, weights: torch.Tensor, x: torch.Tensor, dy: torch.Tensor, dweights: torch.Tensor,
             dx: torch.Tensor | None = None, loss_scale: float | torch.Tensor = 1.0, grad_scale: float = 1.0,
             workspace: torch.Tensor | None = None, n_tiles_dev: torch.Tensor | None = None) -> None:
    """dweights (float32) += grad_scale * dL/dW; dx (float32, if given) = grad_scale * dL/dx.

    ``loss_scale`` may be a one-element float32 GPU tensor (read on the device, no host sync)."""
    n = x.shape[0]
    if n == 0:
        return
    if dy.dtype != torch.float32 or dweights.dtype != torch.float32:
        raise ValueError("dy and dweights must be float32")
    need = workspace_bytes(desc, x.dtype, n)
    if workspace is None or workspace.numel() * workspace.element_size() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=x.device)
    _lib.call("nect_mlp_backward", ctypes.byref(desc.struct), _lib.dtype_code(x.dtype), _lib.ptr(weights),
              x.data_ptr(), n, x.stride(0), None, _lib.ptr(n_tiles_dev), dy.data_ptr(), dy.stride(0),
              1.0 if isinstance(loss_scale, torch.Tensor) else float(loss_scale),
              _lib.ptr(loss_scale) if isinstance(loss_scale, torch.Tensor) else None, float(grad_scale), _lib.ptr(dweights), None if dx is None else dx.data_ptr(),
              0 if dx is None else dx.stride(0), workspace.data_ptr(), workspace.numel() * workspace.element_size(),
              _lib.stream_of(x.device))
