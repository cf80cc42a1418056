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


# ------------------------------------------------------------ wide networks


class GemmMlpDesc:
    """A bias-free MLP of any width, run as plain GEMMs (hipBLASLt through torch.mm) with the activations in
    between -- for networks the fully fused 64-neuron kernels do not cover (tiny-cuda-nn's FullyFusedMLP with
    128 neurons, as NeCT's QuadCubes/HyperCubes presets configure it, or a CutlassMLP of any width).  Same flat
    weight layout as MlpDesc: W0 [width][n_input padded to 16], (H-1) x [width][width], Wout [n_output padded
    to 16][width], row-major."""

    def __init__(self, n_input: int, n_output: int, n_neurons: int, n_hidden_layers: int,
                 activation: str = "ReLU", output_activation: str = "None"):
        if n_neurons < 1 or n_hidden_layers < 1:
            raise ValueError("the network needs at least one hidden layer of at least one neuron")
        activation_code(activation)
        activation_code(output_activation)
        self.n_input = int(n_input)
        self.n_output = int(n_output)
        self.width = int(n_neurons)
        self.n_hidden_layers = int(n_hidden_layers)
        self.n_input_padded = (self.n_input + 15) // 16 * 16
        self.n_output_padded = (self.n_output + 15) // 16 * 16
        self.activation = activation
        self.output_activation = output_activation
        self.n_params = sum(r * c for r, c in self.layer_shapes())

    layer_shapes = MlpDesc.layer_shapes
    split = MlpDesc.split


def _act(name, z: torch.Tensor) -> torch.Tensor:
    from .tcnn import torch_activation

    return torch_activation(name, z)


def _act_grad(name, z: torch.Tensor) -> torch.Tensor:
    """d act / d z at z (float32)."""
    key = "none" if name is None else str(name).casefold()
    if key == "none":
        return torch.ones_like(z)
    if key == "relu":
        return (z > 0).float()
    if key == "leakyrelu":
        return torch.where(z > 0, torch.ones_like(z), torch.full_like(z, 0.01))
    with torch.enable_grad():
        zz = z.detach().requires_grad_(True)
        (g,) = torch.autograd.grad(_act(name, zz).sum(), zz)
    return g


def gemm_forward(desc: GemmMlpDesc, weights: torch.Tensor, x: torch.Tensor):
    """(y [n, n_output] float32, saved pre-activations) of the network on x [n, n_input] in the weights' dtype:
    each layer one GEMM with float32 accumulation, its output rounded to the dtype, the activation applied in
    float32 and rounded again (as a half-precision fused MLP keeps its activations)."""
    dt = weights.dtype
    Ws = desc.split(weights)
    h = x.to(dt)
    if h.shape[1] < desc.n_input_padded:
        h = torch.nn.functional.pad(h, (0, desc.n_input_padded - h.shape[1]))
    zs = []
    for W in Ws[:-1]:
        z = torch.mm(h, W.t().contiguous())
        zs.append(z)
        h = _act(desc.activation, z.float()).to(dt)
    out = torch.mm(h, Ws[-1].t().contiguous())[:, : desc.n_output].float()
    return _act(desc.output_activation, out), (h, zs, out)


def gemm_backward(desc: GemmMlpDesc, weights: torch.Tensor, x: torch.Tensor, saved, dy: torch.Tensor):
    """(dL/dweights [n_params] float32, dL/dx [n, n_input] float32) for dL/dy [n, n_output], in float32 GEMMs."""
    h_last, zs, out = saved
    Ws = [W.float() for W in desc.split(weights)]
    dout = dy.float() * _act_grad(desc.output_activation, out)
    if dout.shape[1] < desc.n_output_padded:
        dout = torch.nn.functional.pad(dout, (0, desc.n_output_padded - dout.shape[1]))
    grads = [None] * len(Ws)
    grads[-1] = torch.mm(dout.t(), h_last.float())
    dh = torch.mm(dout, Ws[-1])
    xp = x.to(weights.dtype)
    if xp.shape[1] < desc.n_input_padded:
        xp = torch.nn.functional.pad(xp, (0, desc.n_input_padded - xp.shape[1]))
    for k in range(desc.n_hidden_layers - 1, -1, -1):
        dz = dh * _act_grad(desc.activation, zs[k].float())
        h_prev = xp if k == 0 else _act(desc.activation, zs[k - 1].float()).to(weights.dtype)
        grads[k] = torch.mm(dz.t(), h_prev.float())
        dh = torch.mm(dz, Ws[k])
    return torch.cat([g.flatten() for g in grads]), dh[:, : desc.n_input]
