"""ORACLE (test infrastructure): one training step of the reference hot path on the CPU.

Restates the batch body of BaseTrainer.fit (nect/trainers/base_trainer.py:
517-653) for a HashGrid model (nect/network/networks.py:28-65):

  1. sample rays (oracle/sampler.c, ct_sampling.cu:606-648) and clamp;
  2. drop all-zero points and compact (base_trainer.py:523-528);
  3. evaluate the model on the kept points (grid indices/weights from
     oracle/hashgrid.py, gather and MLP in float64 torch so autograd yields the
     parameter gradients);
  4. scatter back into [B, S] and form y_pred = sum * (distances / L_max)
     (base_trainer.py:545-554, distances/L_max in float32 as torch does);
  5. L1 / L2 loss (nect/config.py:481-497), backward.

Parameters are the flat [MLP weights | grid table] vector.  Adam is
torch.optim.Adam (library) and is not restated.

Parity: composed of the pinned sampler, hash-grid and MLP oracles plus the
closed-form line integral, which tests/test_oracle_train.py checks against
analytic float64 chord lengths through a sphere.
"""
from __future__ import annotations

import numpy as np
import torch

from . import hashgrid as hg
from . import mlp as om
from .sampler import oracle_sample


def model_forward(spec: hg.GridSpec, mlp_cfg: dict, params: torch.Tensor, x: np.ndarray,
                  round_dtype: torch.dtype | None = None, corners=None) -> torch.Tensor:
    """mu [n] for points x [n, D] (float32 numpy); params float64 torch (may require grad).

    ``round_dtype`` rounds the features and the hidden activations where a
    16-bit implementation stores them (values only, straight-through gradients).
    ``corners`` may pass ``hg.corners(spec, x)`` computed once for repeated calls."""
    n_mlp = mlp_cfg["n_params"]
    W = om.split(params[:n_mlp], mlp_cfg["n_input_padded"], 64, mlp_cfg["n_hidden_layers"], mlp_cfg["n_output_padded"])
    table = params[n_mlp:].view(-1, spec.n_features)
    idx, wts = hg.corners(spec, x) if corners is None else corners
    F = spec.n_features
    feats = []
    for l in range(spec.n_levels):
        rows = torch.from_numpy(spec.offset[l] + idx[:, l, :].astype(np.int64))          # [n, K]
        vals = table[rows.reshape(-1)].view(rows.shape[0], rows.shape[1], F)             # [n, K, F]
        w = torch.from_numpy(wts[:, l, :]).to(params.dtype)
        feats.append((w[:, :, None] * vals).sum(1))
    f = torch.cat(feats, 1)
    if round_dtype is not None:
        f = f + (f.to(round_dtype).to(f.dtype) - f).detach()
    return om.forward(W, f, mlp_cfg["activation"], mlp_cfg["output_activation"], 1, round_dtype=round_dtype)[:, 0]


def integrate(full_mu: torch.Tensor, dist: np.ndarray, l_max: float) -> torch.Tensor:
    """y_pred [B] = sum over samples of mu [B, S] * (distances / L_max), the division in float32
    (base_trainer.py:552-554)."""
    dl = torch.from_numpy((dist / np.float32(l_max)).astype(np.float32)).to(full_mu.dtype)
    return full_mu.sum(1) * dl


class CpuTrainer:
    """The whole training step on the CPU in a chosen float dtype, with torch.optim.Adam: sampler (C
    restatement), compaction, hash grid + MLP (autograd), line integral, L1/L2 loss, backward, Adam --
    BaseTrainer.fit's batch body (nect/trainers/base_trainer.py:517-653) as a CPU program.  Used as the
    measured CPU baseline; its pieces are the pinned functions of this package."""

    def __init__(self, spec: hg.GridSpec, mlp_cfg: dict, params: np.ndarray, lr: float, loss: str = "L2",
                 dtype: torch.dtype = torch.float32):
        self.spec, self.cfg, self.loss, self.dtype = spec, mlp_cfg, loss, dtype
        self.p = torch.tensor(params, dtype=dtype, requires_grad=True)
        self.opt = torch.optim.Adam([self.p], lr=lr)

    def step(self, fields: dict, angle: float, ray_index: np.ndarray, starting_ray_index: int, num_rays: int,
             S: int, step_mm: float, uniform: bool, jitter: np.ndarray, det_offset: np.ndarray, target: np.ndarray,
             l_max: float) -> float:
        pts, dist, _ = oracle_sample(fields, angle, ray_index, num_rays, starting_ray_index, S, step_mm, uniform,
                                     jitter, det_offset)
        flat = pts.reshape(-1, 3)
        keep = ~np.all(flat == 0, axis=1)
        self.opt.zero_grad()
        mu = model_forward(self.spec, self.cfg, self.p, flat[keep])
        full = torch.zeros(num_rays * S, dtype=self.dtype)
        full[torch.from_numpy(keep)] = mu
        y = integrate(full.view(num_rays, S), dist, l_max)
        tgt = torch.from_numpy(np.asarray(target)).to(self.dtype)
        L = (torch.nn.functional.l1_loss(y, tgt) if "l1" in self.loss.casefold()
             else torch.nn.functional.mse_loss(y, tgt))
        L.backward()
        self.opt.step()
        return float(L.detach())


def train_step(fields: dict, angle: float, ray_index: np.ndarray, starting_ray_index: int, num_rays: int, S: int,
               step_mm: float, uniform: bool, jitter: np.ndarray, det_offset: np.ndarray, target: np.ndarray,
               l_max: float, spec: hg.GridSpec, mlp_cfg: dict, params: np.ndarray, loss: str = "L2",
               t: float | None = None, round_dtype: torch.dtype | None = None):
    """Returns dict(loss, y_pred [B], grad [n_params], counts [B], mu_kept [n])."""
    pts, dist, cnt = oracle_sample(fields, angle, ray_index, num_rays, starting_ray_index, S, step_mm, uniform,
                                   jitter, det_offset)
    flat = pts.reshape(-1, 3)
    keep = ~np.all(flat == 0, axis=1)
    x = flat[keep]
    if t is not None:
        x = np.concatenate([x, np.full((x.shape[0], 1), np.float32(t), np.float32)], 1)
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    mu = model_forward(spec, mlp_cfg, p, x, round_dtype=round_dtype)
    full = torch.zeros(num_rays * S, dtype=torch.float64)
    full[torch.from_numpy(keep)] = mu
    y_pred = integrate(full.view(num_rays, S), dist, l_max)
    tgt = torch.from_numpy(np.asarray(target, np.float64))
    key = loss.casefold()
    if "l1" in key:
        L = torch.nn.functional.l1_loss(y_pred, tgt)
    else:
        L = torch.nn.functional.mse_loss(y_pred, tgt)
    L.backward()
    return {"loss": float(L.detach()), "y_pred": y_pred.detach().numpy(), "grad": p.grad.numpy(), "counts": cnt,
            "mu_kept": mu.detach().numpy(), "keep": keep}


def train_step_accumulated(fields: dict, angles, ray_index: np.ndarray, starting_ray_index: int, num_rays: int,
                           S: int, step_mm: float, uniform: bool, jitters, det_offsets, target: np.ndarray,
                           l_max: float, spec: hg.GridSpec, mlp_cfg: dict, params: np.ndarray, loss: str = "L2",
                           t: float | None = None, round_dtype: torch.dtype | None = None):
    """Continuous-scanning step (nect/trainers/continous_scanning_trainer.py:81-167): y_pred of a ray is the
    mean of its line integrals at the K sub-angles (each sampled with its own jitter and detector offset),
    accumulated as sum_k (y_k / K); the loss compares it with the target.  Returns dict(loss, y_pred, grad)."""
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    K = len(angles)
    y_pred = None
    for k in range(K):
        pts, dist, cnt = oracle_sample(fields, angles[k], ray_index, num_rays, starting_ray_index, S, step_mm,
                                       uniform, jitters[k], det_offsets[k])
        flat = pts.reshape(-1, 3)
        keep = ~np.all(flat == 0, axis=1)
        x = flat[keep]
        if t is not None:
            x = np.concatenate([x, np.full((x.shape[0], 1), np.float32(t), np.float32)], 1)
        mu = model_forward(spec, mlp_cfg, p, x, round_dtype=round_dtype)
        full = torch.zeros(num_rays * S, dtype=torch.float64)
        full[torch.from_numpy(keep)] = mu
        yk = integrate(full.view(num_rays, S), dist, l_max) / K
        y_pred = yk if y_pred is None else y_pred + yk
    tgt = torch.from_numpy(np.asarray(target, np.float64))
    if "l1" in loss.casefold():
        L = torch.nn.functional.l1_loss(y_pred, tgt)
    else:
        L = torch.nn.functional.mse_loss(y_pred, tgt)
    L.backward()
    return {"loss": float(L.detach()), "y_pred": y_pred.detach().numpy(), "grad": p.grad.numpy()}
