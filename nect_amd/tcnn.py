"""Drop-in for the tiny-cuda-nn torch bindings that nect/network/networks.py uses.

``import nect_amd.tcnn as tcnn`` gives ``tcnn.Encoding``, ``tcnn.Network`` and
``tcnn.NetworkWithInputEncoding`` with the constructor arguments and config
dictionaries NeCT passes (nect/config.py:86-132, nect/network/networks.py:
36-57, 404-437, 459-496):

  encoding_config  {"otype": "HashGrid", "n_levels", "n_features_per_level",
                    "log2_hashmap_size", "base_resolution", "per_level_scale"}
                   or {"otype": "Composite", "nested": [{"n_dims_to_encode": k,
                   ...}, ...]} with nested "HashGrid" and "Identity" entries
  network_config   {"otype": "FullyFusedMLP" | "CutlassMLP", "n_neurons": 64,
                    "n_hidden_layers", "activation", "output_activation"}

Each module holds one flat float32 ``params`` Parameter (network weights
first, then the encodings' tables, in nesting order) and runs forward and
backward in the gfx950 kernels through torch.autograd.Function.  ``dtype``
selects the compute precision (float16 by default, as tiny-cuda-nn uses on
recent GPUs; float32 and bfloat16 are also available).  Parameter
initialisation is this library's (seeded): tiny-cuda-nn's own streams are not
reproduced.
"""
from __future__ import annotations

import ctypes
import math

import torch
import torch.nn.functional as F

from . import _lib
from . import grid as _grid
from . import mlp as _mlp
from .fused import xavier_init

MAX_GRID_PARTS = 5   # NECT_MAX_GRID_PARTS


def torch_activation(name: str | None, z: torch.Tensor) -> torch.Tensor:
    """Activation definitions shared with the kernels."""
    key = "none" if name is None else str(name).casefold()
    if key == "none":
        return z
    if key == "relu":
        return torch.relu(z)
    if key == "leakyrelu":
        return F.leaky_relu(z, 0.01)
    if key == "exponential":
        return torch.exp(z)
    if key == "sine":
        return torch.sin(z)
    if key == "sigmoid":
        return torch.sigmoid(z)
    if key == "squareplus":
        return 0.5 * (z + torch.sqrt(z * z + 4.0))
    if key == "softplus":
        return F.softplus(z)
    if key == "tanh":
        return torch.tanh(z)
    raise ValueError(f"unsupported activation '{name}'")


# ------------------------------------------------------------- encodings


class _Part:
    """One leaf of an encoding tree: a hash grid or an identity over input columns."""

    def __init__(self, kind: str, col0: int, n_dims: int, out0: int, desc: _grid.GridDesc | None, p0: int):
        self.kind = kind
        self.col0 = col0
        self.n_dims = n_dims
        self.out0 = out0
        self.desc = desc
        self.p0 = p0

    @property
    def n_out(self):
        return self.desc.n_output_dims if self.desc else self.n_dims

    @property
    def n_params(self):
        return self.desc.n_params if self.desc else 0


def _parse_encoding(n_input_dims: int, cfg: dict) -> list[_Part]:
    parts: list[_Part] = []
    col, out, p = 0, 0, 0

    def leaf(c: dict, n_dims: int):
        nonlocal col, out, p
        otype = str(c.get("otype", "")).casefold()
        if otype in ("hashgrid", "grid"):
            if otype == "grid" and str(c.get("type", "Hash")).casefold() != "hash":
                raise ValueError("only hashed grids are supported")
            d = _grid.GridDesc(n_dims, int(c.get("n_levels", 16)), int(c.get("n_features_per_level", 2)),
                               int(c.get("log2_hashmap_size", 19)), int(c.get("base_resolution", 16)),
                               float(c.get("per_level_scale", 2.0)))
            part = _Part("grid", col, n_dims, out, d, p)
        elif otype == "identity":
            part = _Part("identity", col, n_dims, out, None, p)
        else:
            raise ValueError(f"unsupported encoding otype '{c.get('otype')}'")
        parts.append(part)
        col += n_dims
        out += part.n_out
        p += part.n_params

    if str(cfg.get("otype", "")).casefold() == "composite":
        nested = cfg["nested"]
        used = 0
        for i, c in enumerate(nested):
            k = int(c.get("n_dims_to_encode", n_input_dims - used))
            leaf(c, k)
            used += k
        if used != n_input_dims:
            raise ValueError(f"composite encodes {used} dims, the input has {n_input_dims}")
    else:
        leaf(cfg, n_input_dims)
    return parts


def _launch_groups(parts: list[_Part]) -> list[list[_Part]]:
    """The grid parts in launches of nect_grid_*_parts: parts of one (n_dims, n_features), at most
    NECT_MAX_GRID_PARTS per launch."""
    groups: dict = {}
    for p in parts:
        if p.kind == "grid":
            groups.setdefault((p.desc.n_dims, p.desc.n_features), []).append(p)
    out = []
    for g in groups.values():
        for i in range(0, len(g), MAX_GRID_PARTS):
            out.append(g[i:i + MAX_GRID_PARTS])
    return out


def _parts_args(group: list[_Part]):
    n = len(group)
    descs = (_lib.NectGridDesc * n)(*[p.desc.struct for p in group])
    offs = (ctypes.c_int64 * n)(*[p.p0 for p in group])
    cols = (ctypes.c_int32 * n)(*[p.col0 for p in group])
    outs = (ctypes.c_int32 * n)(*[p.out0 for p in group])
    return n, descs, offs, cols, outs


def _row_pitch(width: int, dtype: torch.dtype) -> int:
    """Elements per row of a [n, width] buffer whose rows are written a few bytes per row at a time: widened by 16
    bytes when the rows would be a whole even number of 128-byte lines, whose starts then fall on a fraction of
    the memory channels (measured: the QuadCubes encoder writes 384-column FP16 rows, exactly six lines each,
    1.34x slower than 392-column ones)."""
    esz = torch.empty((), dtype=dtype).element_size()
    row = width * esz
    return width + 16 // esz if row % 256 == 0 else width


def _encode(parts: list[_Part], n_out: int, tables: torch.Tensor, x: torch.Tensor, dtype: torch.dtype,
            width: int | None = None, internal: bool = False) -> torch.Tensor:
    """Features [n, width] of every part (columns past n_out zero): the nested grids in one
    nect_grid_encode_parts launch per (n_dims, n_features), identity parts copied.  ``internal``: rows of
    _row_pitch elements (the features feed the network and are not handed out)."""
    width = n_out if width is None else width
    pitch = _row_pitch(width, dtype) if internal else width
    feats = torch.empty((x.shape[0], pitch), dtype=dtype, device=x.device)[:, :width]
    if width > n_out:
        feats[:, n_out:] = 0
    if x.shape[0] == 0:
        return feats
    for group in _launch_groups(parts):
        n, descs, offs, cols, outs = _parts_args(group)
        _lib.call("nect_grid_encode_parts", n, descs, offs, cols, outs, _lib.dtype_code(dtype), _lib.ptr(tables),
                  _lib.ptr(x), x.shape[0], x.shape[1], feats.data_ptr(), feats.stride(0),
                  _lib.stream_of(x.device))
    for part in parts:
        if part.kind != "grid":
            feats[:, part.out0:part.out0 + part.n_out] = x[:, part.col0:part.col0 + part.n_dims].to(dtype)
    return feats


def _encode_backward(parts: list[_Part], x: torch.Tensor, dfeat: torch.Tensor, grad: torch.Tensor) -> None:
    """grad (laid out as the tables) += the scatter of dfeat [n, n_out] (float, unit column stride) for every
    grid part."""
    if x.shape[0] == 0:
        return
    if dfeat.stride(1) != 1:
        dfeat = dfeat.contiguous()
    for group in _launch_groups(parts):
        n, descs, offs, cols, outs = _parts_args(group)
        _lib.call("nect_grid_backward_parts", n, descs, offs, cols, outs, _lib.ptr(x), x.shape[0], x.shape[1],
                  dfeat.data_ptr(), dfeat.stride(0), _lib.ptr(grad), _lib.stream_of(x.device))


def _check_input(x: torch.Tensor, n: int) -> torch.Tensor:
    if not x.is_cuda:
        raise ValueError("nect_amd modules run on the GPU only")
    if x.dim() != 2 or x.shape[1] != n:
        raise ValueError(f"expected input [N, {n}], got {tuple(x.shape)}")
    return x.float().contiguous()


def _loss_scale(dy: torch.Tensor) -> torch.Tensor:
    """Power-of-two scale bringing max|dy| near 1 before the low-precision products (device scalar)."""
    m = dy.abs().amax().clamp_min(1e-30)
    return torch.exp2(-torch.floor(torch.log2(m))).reshape(1).float()


class _EncodingFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, params, module):
        tables = params.to(module.dtype)
        feats = _encode(module.parts, module.n_output_dims, tables, x, module.dtype)
        ctx.save_for_backward(x)
        ctx.module = module
        return feats

    @staticmethod
    def backward(ctx, dfeat):
        (x,) = ctx.saved_tensors
        grad = torch.zeros(ctx.module.n_params, dtype=torch.float32, device=x.device)
        _encode_backward(ctx.module.parts, x, dfeat.float().contiguous(), grad)
        return None, grad, None


class Encoding(torch.nn.Module):
    """tcnn.Encoding(n_input_dims, encoding_config, seed=1337, dtype=None)."""

    def __init__(self, n_input_dims: int, encoding_config: dict, seed: int = 1337, dtype: torch.dtype | None = None):
        super().__init__()
        self.n_input_dims = int(n_input_dims)
        self.encoding_config = encoding_config
        self.dtype = dtype or torch.float16
        self.parts = _parse_encoding(self.n_input_dims, encoding_config)
        self.n_output_dims = sum(p.n_out for p in self.parts)
        self.n_params = sum(p.n_params for p in self.parts)
        g = torch.Generator().manual_seed(seed)
        self.params = torch.nn.Parameter((torch.rand(self.n_params, generator=g) * 2 - 1) * 1e-4)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        x = _check_input(x, self.n_input_dims)
        return _EncodingFunction.apply(x, self.params, self)


# --------------------------------------------------------------- networks


def _mlp_desc(n_input: int, n_output: int, cfg: dict, dtype: torch.dtype = torch.float16):
    """The fully fused kernels' MlpDesc for 64 neurons, 1-8 hidden layers and up to 128 (padded) inputs (float32:
    up to 3 hidden layers and 64 inputs, what its twice-as-large LDS image leaves the backward); otherwise the
    GEMM network (GemmMlpDesc)."""
    otype = str(cfg.get("otype", "FullyFusedMLP")).casefold()
    if otype not in ("fullyfusedmlp", "cutlassmlp", "mlp"):
        raise ValueError(f"unsupported network otype '{cfg.get('otype')}'")
    width, hidden = int(cfg.get("n_neurons", 64)), int(cfg.get("n_hidden_layers", 2))
    act, out_act = cfg.get("activation", "ReLU"), cfg.get("output_activation", "None")
    padded = (n_input + 15) // 16 * 16
    max_hidden, max_padded = (3, 64) if dtype == torch.float32 else (8, 128)
    if width == 64 and 1 <= hidden <= max_hidden and padded <= max_padded and n_output <= 16:
        return _mlp.MlpDesc(n_input, n_output, width, hidden, act, out_act)
    return _mlp.GemmMlpDesc(n_input, n_output, width, hidden, act, out_act)


def _net_forward(desc, w: torch.Tensor, x: torch.Tensor):
    """(y float32, state for _net_backward)."""
    if isinstance(desc, _mlp.GemmMlpDesc):
        return _mlp.gemm_forward(desc, w, x)
    return _mlp.forward(desc, w, x), None


def _net_backward(desc, w: torch.Tensor, x: torch.Tensor, saved, dy: torch.Tensor, dw: torch.Tensor,
                  keep_scaled: bool = False):
    """dw += dL/dW; returns (dx, dx_scale): dL/dx [n, n_input] float32 and None (GEMM networks: the leading
    columns of a wider tensor, rows of unit column stride), or with ``keep_scaled`` a GEMM network's 16-bit
    scaled result and the float32 [1] device factor that unscales it (see mlp.gemm_backward)."""
    if isinstance(desc, _mlp.GemmMlpDesc):
        g, d, d_scale = _mlp.gemm_backward(desc, w, x, saved, dy, keep_scaled)
        dw += g
        return d[:, : desc.n_input], d_scale
    dx = torch.empty((x.shape[0], desc.n_input), dtype=torch.float32, device=dy.device)
    _mlp.backward(desc, w, x, dy, dw, dx, loss_scale=_loss_scale(dy))
    return dx, None


class _NetworkFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, params, module):
        w = params.to(module.dtype)
        xd = x.to(module.dtype)
        y, saved = _net_forward(module.desc, w, xd)
        ctx.save_for_backward(xd, w)
        ctx.saved_state = saved
        ctx.module = module
        return y

    @staticmethod
    def backward(ctx, dy):
        xd, w = ctx.saved_tensors
        m = ctx.module
        dy = dy.float().contiguous()
        dw = torch.zeros(m.desc.n_params, dtype=torch.float32, device=dy.device)
        dx, _ = _net_backward(m.desc, w, xd, ctx.saved_state, dy, dw)
        return dx, dw, None


class Network(torch.nn.Module):
    """tcnn.Network(n_input_dims, n_output_dims, network_config, seed=1337)."""

    def __init__(self, n_input_dims: int, n_output_dims: int, network_config: dict, seed: int = 1337,
                 dtype: torch.dtype | None = None):
        super().__init__()
        self.n_input_dims = int(n_input_dims)
        self.n_output_dims = int(n_output_dims)
        self.network_config = network_config
        self.dtype = dtype or torch.float16
        self.desc = _mlp_desc(self.n_input_dims, self.n_output_dims, network_config, self.dtype)
        self.params = torch.nn.Parameter(xavier_init(self.desc, torch.Generator().manual_seed(seed)))

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        x = _check_input(x, self.n_input_dims)
        return _NetworkFunction.apply(x, self.params, self)


class _NetworkWithEncodingFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, params, module):
        n_net = module.desc.n_params
        p = module.compute_params
        if p is None:
            p = params.to(module.dtype)
        elif p.dtype != module.dtype or p.shape != params.shape or p.device != params.device:
            raise ValueError("compute_params must be the parameters' size, device and compute dtype")
        w, tables = p[:n_net], p[n_net:]
        # written at the network's padded input width (GEMM networks) or a multiple of 8 (fused kernels)
        if isinstance(module.desc, _mlp.GemmMlpDesc):
            width = module.desc.n_input_padded
        else:
            width = (module.n_encoded + 7) // 8 * 8
        feats = _encode(module.parts, module.n_encoded, tables, x, module.dtype, width, internal=True)
        y, saved = _net_forward(module.desc, w, feats)
        ctx.save_for_backward(x, w, feats)
        ctx.saved_state = saved
        ctx.module = module
        return y

    @staticmethod
    def backward(ctx, dy):
        x, w, feats = ctx.saved_tensors
        m = ctx.module
        dy = dy.float().contiguous()
        buf = m.grad_buffer
        n_all = m.desc.n_params + m.n_enc_params
        if buf is not None and (buf.dtype != torch.float32 or buf.numel() != n_all or not buf.is_contiguous()
                                or buf.device != dy.device):
            raise ValueError(f"grad_buffer must be a contiguous float32 tensor of {n_all} elements on {dy.device}")
        grad = buf if buf is not None else torch.zeros(n_all, dtype=torch.float32, device=dy.device)
        hook = m.encoding_backward
        dfeat, dfeat_scale = _net_backward(m.desc, w, feats, ctx.saved_state, dy, grad[: m.desc.n_params],
                                           keep_scaled=hook is not None)
        if hook is not None:
            # a caller that knows where the rows came from scatters the table gradient itself
            hook(dfeat, grad[m.desc.n_params:], dfeat_scale)
        else:
            _encode_backward(m.parts, x, dfeat, grad[m.desc.n_params:])
        return None, (None if buf is not None else grad), None


class NetworkWithInputEncoding(torch.nn.Module):
    """tcnn.NetworkWithInputEncoding(n_input_dims, n_output_dims, encoding_config, network_config, seed=1337)."""

    def __init__(self, n_input_dims: int, n_output_dims: int, encoding_config: dict, network_config: dict,
                 seed: int = 1337, dtype: torch.dtype | None = None):
        super().__init__()
        self.n_input_dims = int(n_input_dims)
        self.n_output_dims = int(n_output_dims)
        self.encoding_config = encoding_config
        self.network_config = network_config
        self.dtype = dtype or torch.float16
        self.parts = _parse_encoding(self.n_input_dims, encoding_config)
        self.n_encoded = sum(p.n_out for p in self.parts)
        self.n_enc_params = sum(p.n_params for p in self.parts)
        self.desc = _mlp_desc(self.n_encoded, self.n_output_dims, network_config, self.dtype)
        g = torch.Generator().manual_seed(seed)
        w = xavier_init(self.desc, g)
        tables = (torch.rand(self.n_enc_params, generator=g) * 2 - 1) * 1e-4
        self.params = torch.nn.Parameter(torch.cat([w, tables]))
        # optional callable(dfeat [n, n_encoded] of unit column stride, table gradient view, dfeat_scale)
        # replacing the encoding's own backward (the training step sets it to scatter over ray groups):
        # dL/dfeatures = dfeat (float32) when dfeat_scale is None, else dfeat (the network's 16-bit loss-scaled
        # result) times dfeat_scale (a float32 [1] device tensor)
        self.encoding_backward = None
        # optional float32 tensor of the parameters' size: the backward adds the parameter gradient into it and
        # leaves params.grad alone (an optimiser that owns the buffer reads and clears it, and a model of a
        # billion parameters is spared a fresh zero-filled gradient per step)
        self.grad_buffer = None
        # optional tensor of the parameters' size in the compute dtype, equal to params.to(dtype): the forward
        # uses it instead of casting the parameters (a training step whose optimiser writes that copy as it
        # updates the parameters sets it for the step's own forward and backward)
        self.compute_params = None

    @property
    def n_params(self) -> int:
        return self.params.numel()

    def grid_parts(self):
        return [p for p in self.parts if p.kind == "grid"]

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        x = _check_input(x, self.n_input_dims)
        return _NetworkWithEncodingFunction.apply(x, self.params, self)


__all__ = ["Encoding", "Network", "NetworkWithInputEncoding", "torch_activation"]
