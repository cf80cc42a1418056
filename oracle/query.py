"""ORACLE (test infrastructure): the export queries of the reference on the CPU.

Restates how the reference builds the voxel grid of an export and turns the
network output into voxel values:

  static export   nect/static_export.py:53-141: ROI/binning sizes (:53-88),
                  coordinates by torch.linspace on the CPU (:109-111), voxel
                  v -> (v // (y*x), (v % (y*x)) // x, v % x) (:121-128),
                  rescale on CPU float32 tensors (:138-140);
  dynamic export  nect/dynamic_export.py:473-541 (export_volumes_at_timesteps):
                  the same sizes (:485-516), timesteps divided by the number
                  of angles when any exceeds 1 (:517-518), coordinates by
                  torch.linspace on the GPU (:521-531), rescale on GPU float32
                  tensors (:535-537);
  create_volume   nect/trainers/base_trainer.py:402-450: voxel centres
                  linspace(0.5/n, 1 - 0.5/n, n) on the GPU, y/x sliced by
                  sample_outside (:408-436), rescale on the network's
                  (half) output tensor on the GPU (:442-444).

Rules restated here, and what pins each (tests/test_oracle_query.py on the CPU,
tests/test_query_gpu.py on the device):

  linspace_device  torch's GPU linspace (ATen RangeFactories: float32
                   step = (b - a) / (n - 1); a + step*i below n // 2 and
                   b - step*(n - 1 - i) from there, each a fused multiply-add
                   as the device compiler contracts it).  Pinned by closed
                   forms (endpoints, exactly representable grids, distance to
                   the exact rational grid) and, on the GPU, bit-exactly
                   against torch.linspace(device="cuda").
  linspace_cpu     torch.linspace(..., device="cpu") itself: the call the
                   static export makes (torch is the library the reference
                   runs; its CPU kernel is not restated).
  rescale          ((mu / l_max) * span) + lo, each operation rounded:
                   "divide" (float32 true division: torch on CPU tensors),
                   "reciprocal" (mu * fl(1/l_max): torch's GPU division by a
                   scalar), "reciprocal_f16" (the same with every result
                   rounded to half: a half tensor on the GPU).  Pinned against
                   torch's CPU operations here and its GPU operations on the
                   device.
"""
from __future__ import annotations

import numpy as np
import torch

from .hashgrid import fma32

RESCALE = {"divide": 0, "reciprocal": 1, "reciprocal_f16": 2}


def linspace_device(a: float, b: float, n: int) -> np.ndarray:
    """torch.linspace(a, b, n) in float32 as torch computes it on the GPU."""
    a32, b32 = np.float32(a), np.float32(b)
    if n == 1:
        return np.array([a32], np.float32)
    step = np.float32((b32 - a32) / np.float32(n - 1))
    i = np.arange(n, dtype=np.int64)
    half = n // 2
    lo = fma32(step, i[:half].astype(np.float32), a32)
    hi = fma32(-step, (n - 1 - i[half:]).astype(np.float32), b32)
    return np.concatenate([lo, hi]).astype(np.float32)


def linspace_cpu(a: float, b: float, n: int) -> np.ndarray:
    """torch.linspace(a, b, n, dtype=float32, device="cpu"), the static export's call."""
    return torch.linspace(a, b, steps=n, dtype=torch.float32, device="cpu").numpy()


def export_extent(nVoxel, binning: int = 1, ROIx=None, ROIy=None, ROIz=None, sample_outside: int = 0):
    """((z_h, y_w, x_w), (start_z, end_z), (start_y, end_y), (start_x, end_x)) of an export.

    nect/static_export.py:53-88 and nect/dynamic_export.py:485-516 (identical): the full grid takes
    z_h = nVoxel[0] // binning and y_w = x_w = nVoxel[1] // binning (the width stands for both in-plane
    sizes); an ROI [i0, i1) of the padded grid (sample_outside voxels on each side in y and x) spans
    (i0 - rm) / n_padded .. (i1 - rm) / n_padded with (i1 - i0) // binning voxels (z: no padding)."""
    height, width = int(nVoxel[0]), int(nVoxel[1])
    z_h = height // binning
    y_w = width // binning
    x_w = width // binning
    rm = int(sample_outside)
    n_pad = [int(nVoxel[0]), int(nVoxel[1]) + 2 * rm, int(nVoxel[2]) + 2 * rm]
    sx, ex = 0, 1
    if ROIx is not None:
        sx = (ROIx[0] - rm) / n_pad[2]
        ex = (ROIx[1] - rm) / n_pad[2]
        x_w = (ROIx[1] - ROIx[0]) // binning
    sy, ey = 0, 1
    if ROIy is not None:
        sy = (ROIy[0] - rm) / n_pad[1]
        ey = (ROIy[1] - rm) / n_pad[1]
        y_w = (ROIy[1] - ROIy[0]) // binning
    sz, ez = 0, 1
    if ROIz is not None:
        sz = ROIz[0] / n_pad[0]
        ez = ROIz[1] / n_pad[0]
        z_h = (ROIz[1] - ROIz[0]) // binning
    return (z_h, y_w, x_w), (sz, ez), (sy, ey), (sx, ex)


def static_export_axes(nVoxel, binning: int = 1, ROIx=None, ROIy=None, ROIz=None, sample_outside: int = 0):
    """(z, y, x) coordinate arrays of export_volume (CPU linspace)."""
    (z_h, y_w, x_w), zr, yr, xr = export_extent(nVoxel, binning, ROIx, ROIy, ROIz, sample_outside)
    return linspace_cpu(*zr, z_h), linspace_cpu(*yr, y_w), linspace_cpu(*xr, x_w)


def dynamic_export_axes(nVoxel, binning: int = 1, ROIx=None, ROIy=None, ROIz=None, sample_outside: int = 0):
    """(z, y, x) coordinate arrays of export_volumes_at_timesteps (GPU linspace)."""
    (z_h, y_w, x_w), zr, yr, xr = export_extent(nVoxel, binning, ROIx, ROIy, ROIz, sample_outside)
    return linspace_device(*zr, z_h), linspace_device(*yr, y_w), linspace_device(*xr, x_w)


def export_timesteps(timesteps, n_angles: int) -> list[float]:
    """nect/dynamic_export.py:517-518: timesteps above 1 are projection numbers, divided by the angle count."""
    ts = list(timesteps)
    if max(ts) > 1:
        ts = [t / n_angles for t in ts]
    return ts


def create_volume_axes(nVoxel, sample_outside: int = 0):
    """(z, y, x) voxel-centre coordinates of create_volume (GPU linspace, y/x sliced by sample_outside)."""
    rm = int(sample_outside)
    size = [int(v) for v in nVoxel]
    padded = [size[0], size[1] + 2 * rm, size[2] + 2 * rm]
    z = linspace_device(0.5 / size[0], 1 - 0.5 / size[0], size[0])
    y = linspace_device(0.5 / padded[1], 1 - 0.5 / padded[1], padded[1])
    x = linspace_device(0.5 / padded[2], 1 - 0.5 / padded[2], padded[2])
    if rm > 0:
        y, x = y[rm:-rm], x[rm:-rm]
    return z, y, x


def grid_points(axes, t: float | None = None) -> np.ndarray:
    """Points [nz*ny*nx, 3 or 4] (z, y, x[, t]) in the export's voxel order."""
    z, y, x = axes
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    cols = [zz.ravel(), yy.ravel(), xx.ravel()]
    if t is not None:
        cols.append(np.full(zz.size, np.float32(t), np.float32))
    return np.stack(cols, 1).astype(np.float32)


def rescale(mu: np.ndarray, l_max: float, span: float, lo: float, mode: str = "divide",
            mu_dtype: str = "float32") -> np.ndarray:
    """((mu / l_max) * span) + lo as the reference's export computes it (see the module docstring);
    mu_dtype "float16" first rounds mu to half (the output of a half-precision network)."""
    mu = np.asarray(mu, np.float32)
    if mu_dtype == "float16":
        mu = mu.astype(np.float16).astype(np.float32)
    L, S, O = np.float32(l_max), np.float32(span), np.float32(lo)
    if mode == "divide":
        return ((mu / L) * S + O).astype(np.float32)
    inv = np.float32(np.float32(1.0) / L)
    if mode == "reciprocal":
        return (((mu * inv).astype(np.float32) * S).astype(np.float32) + O).astype(np.float32)
    if mode == "reciprocal_f16":
        h = mu.astype(np.float16)
        h = (h.astype(np.float32) * inv).astype(np.float16)
        h = (h.astype(np.float32) * S).astype(np.float16)
        h = (h.astype(np.float32) + O).astype(np.float16)
        return h.astype(np.float32)
    raise ValueError(mode)
