"""ORACLE (test infrastructure): bias-free MLP in float64 (torch CPU).

Restates tiny-cuda-nn's FullyFusedMLP as NeCT configures it
(nect/config.py:125-132, activation names nect/config.py:704-714): no biases,
``n_hidden_layers`` hidden layers, activation after every hidden layer and
``output_activation`` on the output.  Weights are the flat vector split into
row-major [out][in] matrices: W0 [width][n_in_pad], (H-1) x [width][width],
Wout [n_out_pad][width]; padded inputs are zero, padded outputs are dropped.

Activation definitions (the build's specification; tiny-cuda-nn's internal
scaling of softplus/squareplus is unpinned): ReLU, LeakyReLU (slope 0.01),
Exponential, Sine, Sigmoid, Squareplus 0.5*(z + sqrt(z^2 + 4)),
Softplus log(1 + e^z) (torch's, threshold 20), Tanh.

Parity pins: this is the textbook definition (F.linear chain); its
gradients come from torch.autograd, checked by gradcheck in
tests/test_oracle_mlp.py.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F


def activation(name: str | None, z: torch.Tensor) -> torch.Tensor:
    key = "none" if name is None else name.casefold()
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
    raise ValueError(name)


def split(flat: torch.Tensor, n_in_pad: int, width: int, n_hidden: int, n_out_pad: int) -> list[torch.Tensor]:
    shapes = [(width, n_in_pad)] + [(width, width)] * (n_hidden - 1) + [(n_out_pad, width)]
    out, o = [], 0
    for r, c in shapes:
        out.append(flat[o:o + r * c].view(r, c))
        o += r * c
    return out


def forward(weights: list[torch.Tensor], x: torch.Tensor, act: str, out_act: str, n_output: int,
            round_dtype: torch.dtype | None = None) -> torch.Tensor:
    """x [n, n_input] -> y [n, n_output]; inputs are zero-padded to W0's width.

    ``round_dtype`` rounds each hidden activation that feeds another hidden layer
    to that dtype (value only; gradients pass straight through), which is where
    a 16-bit matrix-core implementation rounds them; the activation feeding the
    output layer is used unrounded."""
    n_in_pad = weights[0].shape[1]
    if x.shape[1] < n_in_pad:
        x = F.pad(x, (0, n_in_pad - x.shape[1]))
    h = x
    hidden = weights[:-1]
    for k, W in enumerate(hidden):
        h = activation(act, F.linear(h, W))
        if round_dtype is not None and k < len(hidden) - 1:
            h = h + (h.to(round_dtype).to(h.dtype) - h).detach()
    z = F.linear(h, weights[-1])[:, :n_output]
    return activation(out_act, z)
