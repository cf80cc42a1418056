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


_PWL = {"none": None, "relu": 0.0, "leakyrelu": 0.01}


def _pwl_slope(name):
    """The negative-side slope of a piecewise-linear activation with act(0) = 0 (None: not piecewise linear)."""
    key = "none" if name is None else str(name).casefold()
    return (1.0 if key == "none" else _PWL[key]) if key in _PWL else None


# Whether 16-bit layers whose shapes nect_gemm_act covers take it (else torch.mm and torch activations).
FUSED_LAYERS = True


def _fused_ok(dt: torch.dtype, k: int, m: int, slope) -> bool:
    return (FUSED_LAYERS and slope is not None and dt in (torch.float16, torch.bfloat16) and m in (128, 384)
            and k % 16 == 0 and 16 <= k <= 384 and (m + 64) * (k + 8) * 2 <= 160 * 1024)


def gemm_act(a: torch.Tensor, bt: torch.Tensor, epilogue: int = 0, slope: float = 1.0,
             h: torch.Tensor | None = None) -> torch.Tensor:
    """epi(a @ bt^T) [n, m] in a's dtype (nect_gemm_act): a [n, k] and bt [m, k] 16-bit with unit column stride
    and 16-byte aligned rows; epilogue 1 applies z > 0 ? z : slope * z, epilogue 2 masks with the saved
    activations h [n, m] (h > 0 ? z : slope * z)."""
    n, k = a.shape
    m = bt.shape[0]
    out = torch.empty((n, m), dtype=a.dtype, device=a.device)
    _lib.call("nect_gemm_act", _lib.dtype_code(a.dtype), a.data_ptr(), a.stride(0), bt.data_ptr(), bt.stride(0), n,
              k, m, int(epilogue), float(slope), None if h is None else h.data_ptr(),
              0 if h is None else h.stride(0), out.data_ptr(), out.stride(0), _lib.stream_of(a.device))
    return out


def gemm_forward(desc: GemmMlpDesc, weights: torch.Tensor, x: torch.Tensor):
    """(y [n, n_output] float32, saved state) of the network on x [n, n_input] in the weights' dtype: each
    layer one GEMM with float32 accumulation, its output rounded to the dtype, the activation computed in
    float32 and rounded once more (as a half-precision fused MLP keeps its activations; torch evaluates
    activations of 16-bit tensors in float32 and rounds, so they are applied to the dtype's tensors directly)."""
    dt = weights.dtype
    Ws = desc.split(weights)
    h = x.to(dt)
    if h.shape[1] < desc.n_input_padded:
        h = torch.nn.functional.pad(h, (0, desc.n_input_padded - h.shape[1]))
    pwl = _pwl_slope(desc.activation) is not None
    one_op = str(desc.activation).casefold() != "squareplus"   # several torch ops: evaluate it in float32
    slope = _pwl_slope(desc.activation)
    zs, hs = [], []   # the input is not kept: the backward pads x again
    for W in Ws[:-1]:
        if _fused_ok(dt, h.shape[1], W.shape[0], slope) and h.stride(1) == 1:
            # product and activation in one kernel, rounded once (as the fused 64-neuron kernels do)
            h = gemm_act(h, W, 1, slope)
            zs.append(None)
            hs.append(h)
            continue
        z = torch.mm(h, W.t().contiguous())
        h = _act(desc.activation, z) if one_op else _act(desc.activation, z.float())
        if h.dtype != dt:
            h = h.to(dt)
        zs.append(None if pwl else z)
        hs.append(h)
    out = torch.mm(h, Ws[-1].t().contiguous())[:, : desc.n_output].float()
    y = _act(desc.output_activation, out)
    # the state must not hold the returned tensor itself (an identity output activation returns `out`): the
    # autograd node keeps the state, the output keeps its node, and that cycle would hold every layer's
    # activations until Python's collector ran.  act'(z) of an identity is 1 and needs no z.
    return y, (hs, zs, None if y is out else out)


def wgrad(dz: torch.Tensor, h: torch.Tensor) -> torch.Tensor:
    """dz^T @ h [dz columns, h columns] float32 for rows of samples: float32 GEMM for float32 operands, else
    nect_wgrad (the samples split across the device, float32 accumulation)."""
    if dz.dtype == torch.float32:
        return torch.mm(dz.t(), h)
    if dz.dtype != h.dtype or dz.stride(1) != 1 or h.stride(1) != 1 or not dz.is_cuda:
        raise ValueError("wgrad takes two GPU matrices of one 16-bit dtype with unit column stride")
    out = torch.zeros((dz.shape[1], h.shape[1]), dtype=torch.float32, device=dz.device)
    _lib.call("nect_wgrad", _lib.dtype_code(dz.dtype), dz.data_ptr(), dz.stride(0), h.data_ptr(), h.stride(0),
              dz.shape[0], dz.shape[1], h.shape[1], 1.0, out.data_ptr(), out.stride(0), _lib.stream_of(dz.device))
    return out


def gemm_backward(desc: GemmMlpDesc, weights: torch.Tensor, x: torch.Tensor, saved, dy: torch.Tensor,
                  keep_scaled: bool = False):
    """(dL/dweights [n_params] float32, dL/dx [n, >= n_input]: the first n_input columns, dx_scale) for dL/dy
    [n, n_output].  dL/dx is float32 and dx_scale None, except with ``keep_scaled`` and 16-bit weights: then
    dL/dx is the backward's own 16-bit result, still multiplied by the loss scale, and dx_scale the reciprocal
    of that power of two (a float32 [1] device tensor) -- dL/dx = dx * dx_scale exactly, for a consumer that
    applies it as it reads (nect_grid_backward_ray_groups_cols).

    16-bit weights: dL/dz is scaled by a power of two of max|dL/dy| (on the device) and kept in the dtype,
    dL/dx and the layer inputs' gradients are GEMMs in the dtype (K = width), the weight gradients GEMMs over
    the samples (nect_wgrad) with float32 accumulation and result, unscaled exactly; float32 weights: float32
    GEMMs.  For piecewise-linear activations act'(z) follows from the sign of the saved act(z)."""
    hs, zs, out = saved
    dt = weights.dtype
    Ws = desc.split(weights)
    slope = _pwl_slope(desc.activation)
    if dt == torch.float32:
        scale = None
        dout = dy.float() if out is None else dy.float() * _act_grad(desc.output_activation, out)
    else:
        from .tcnn import _loss_scale

        scale = _loss_scale(dy)
        dout = (dy.float() * scale if out is None else dy.float() * _act_grad(desc.output_activation, out) * scale).to(dt)
    if dout.shape[1] < desc.n_output_padded:
        dout = torch.nn.functional.pad(dout, (0, desc.n_output_padded - dout.shape[1]))
    grads = [None] * len(Ws)
    grads[-1] = wgrad(dout, hs[-1])
    # dz of the layer below, masked in the product's epilogue when nect_gemm_act covers it
    fused_dz = None
    Wt = Ws[-1].t().contiguous()
    if _fused_ok(dt, dout.shape[1], Wt.shape[0], slope):
        fused_dz = gemm_act(dout, Wt, 2, slope, hs[-1])
        dh = None
    else:
        dh = torch.mm(dout, Ws[-1])
    for k in range(desc.n_hidden_layers - 1, -1, -1):
        if fused_dz is not None:
            dz = fused_dz
        elif slope is not None:
            # one fused kernel each, from the saved act(z) (act(z) > 0 exactly where z > 0)
            if slope == 1.0:
                dz = dh
            elif slope == 0.0:
                dz = torch.ops.aten.threshold_backward(dh, hs[k], 0)
            else:
                dz = torch.ops.aten.leaky_relu_backward(dh, hs[k], slope, True)
        else:
            dz = (dh.float() * _act_grad(desc.activation, zs[k].float())).to(dh.dtype)
        if k > 0:
            h_prev = hs[k - 1]
        else:
            h_prev = x.to(dt)
            if h_prev.shape[1] < desc.n_input_padded:
                h_prev = torch.nn.functional.pad(h_prev, (0, desc.n_input_padded - h_prev.shape[1]))
        grads[k] = wgrad(dz, h_prev)
        fused_dz = None
        Wt = Ws[k].t().contiguous()
        if k > 0 and _fused_ok(dt, dz.shape[1], Wt.shape[0], slope):
            fused_dz = gemm_act(dz, Wt, 2, slope, hs[k - 1])
        elif k == 0 and keep_scaled and scale is not None and _fused_ok(dt, dz.shape[1], Wt.shape[0], 1.0):
            dh = gemm_act(dz, Wt)          # dL/dx, still scaled, for a consumer that unscales it
        else:
            dh = torch.mm(dz, Ws[k])
    g = torch.cat([w.flatten() for w in grads])
    if scale is not None and keep_scaled:
        return g / scale, dh, torch.reciprocal(scale)
    if scale is not None:
        from .tcnn import _row_pitch

        g = g / scale
        # 16-bit / float32 [1]: float32 in one pass, exact (a power of two); rows padded as the features' are
        dx = torch.empty((dh.shape[0], _row_pitch(dh.shape[1], torch.float32)), dtype=torch.float32,
                         device=dh.device)[:, : dh.shape[1]]
        dh = torch.div(dh, scale, out=dx)
    return g, dh, None
