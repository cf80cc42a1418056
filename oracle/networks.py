"""ORACLE (test infrastructure): NeCT's composite models in float64 (torch CPU, differentiable).

Restates the models of nect/network/networks.py that the fused step does not cover:

  * input assembly: QuadCubes (:437-449) feeds [zyx, yxt, xzt, zyt(, zyxt)] -- yxt = (y, x, t), xzt = (x, z, t),
    zyt = (z, y, t), t the per-view time in every row; HyperCubes (:496-508) adds zyxt for its 4-D grid (and
    once more for the identity); HashGrid with include_identity (:36-49, :61-64) feeds [zyx, zyx];
  * tiny-cuda-nn's Composite encoding (nested encodings over consecutive input columns, outputs concatenated
    in order; Identity passes its columns through), as the build's specification (tiny-cuda-nn is not part of
    the reference tree);
  * the bias-free MLP of any width (oracle/mlp.py), weights first in the flat parameter vector, then each
    grid's table in nesting order.

Pins (tests/test_oracle_composite.py): a Composite with one nested grid over all columns equals the plain
grid model of oracle/train.py; identity columns come out unchanged; the input assembly against hand-built
rows; the gradient against central differences.
"""
from __future__ import annotations

import numpy as np
import torch

from . import hashgrid as hg
from . import mlp as om


def quad_inputs(zyx: np.ndarray, t: float, identity: bool = True) -> np.ndarray:
    z, y, x = zyx[:, 0:1], zyx[:, 1:2], zyx[:, 2:3]
    tt = np.full_like(z, np.float32(t))
    cols = [zyx, np.concatenate([y, x, tt], 1), np.concatenate([x, z, tt], 1), np.concatenate([z, y, tt], 1)]
    if identity:
        cols.append(np.concatenate([zyx, tt], 1))
    return np.concatenate(cols, 1).astype(np.float32)


def hyper_inputs(zyx: np.ndarray, t: float, identity: bool = True) -> np.ndarray:
    zyxt = np.concatenate([zyx, np.full_like(zyx[:, :1], np.float32(t))], 1)
    base = quad_inputs(zyx, t, identity=False)
    cols = [base, zyxt] + ([zyxt] if identity else [])
    return np.concatenate(cols, 1).astype(np.float32)


def grid_features(spec: hg.GridSpec, table: torch.Tensor, x: np.ndarray) -> torch.Tensor:
    """Features [n, L*F] of one grid (differentiable in table, float64)."""
    idx, wts = hg.corners(spec, x)
    F = spec.n_features
    tab = table.view(-1, F)
    feats = []
    for l in range(spec.n_levels):
        rows = torch.from_numpy(spec.offset[l] + idx[:, l, :].astype(np.int64))
        vals = tab[rows.reshape(-1)].view(rows.shape[0], rows.shape[1], F)
        feats.append((torch.from_numpy(wts[:, l, :]).to(table.dtype)[:, :, None] * vals).sum(1))
    return torch.cat(feats, 1)


def composite_forward(parts, mlp_cfg: dict, params: torch.Tensor, x: np.ndarray,
                      round_dtype: torch.dtype | None = None) -> torch.Tensor:
    """mu [n] of a Composite-encoded model.  parts: list of ("grid", GridSpec) or ("identity", n_dims) in nesting
    order, consuming consecutive columns of x [n, n_in] (float32); mlp_cfg: n_params, n_input_padded, width,
    n_hidden_layers, n_output_padded, activation, output_activation.  ``round_dtype`` rounds the features and
    hidden activations where a 16-bit implementation stores them (values only)."""
    n_mlp = mlp_cfg["n_params"]
    W = om.split(params[:n_mlp], mlp_cfg["n_input_padded"], mlp_cfg.get("width", 64), mlp_cfg["n_hidden_layers"],
                 mlp_cfg["n_output_padded"])
    col, p0, feats = 0, n_mlp, []
    for kind, arg in parts:
        if kind == "grid":
            spec = arg
            d = spec.n_dims
            feats.append(grid_features(spec, params[p0:p0 + spec.n_params], np.ascontiguousarray(x[:, col:col + d])))
            p0 += spec.n_params
        else:
            d = arg
            feats.append(torch.from_numpy(x[:, col:col + d].astype(np.float64)).to(params.dtype))
        col += d
    f = torch.cat(feats, 1)
    if round_dtype is not None:
        f = f + (f.to(round_dtype).to(f.dtype) - f).detach()
    return om.forward(W, f, mlp_cfg["activation"], mlp_cfg["output_activation"], 1, round_dtype=round_dtype)[:, 0]
