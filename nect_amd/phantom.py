"""Seeded synthetic phantoms and their analytic projections (data generation, not the hot path).

A phantom is a sum of ellipsoids (attenuation per mm, additive) in the volume
frame of the sampler: millimetres, volume centred on the origin, x along the
source-detector axis at angle 0, z along detector rows.  Projections are exact
line integrals computed in float64 with torch (on the GPU when the rays are
there): the chord of each ellipsoid follows from a quadratic in its unit-sphere
frame.  The ray of detector pixel (row, col) joins the same source and
detector-pixel-centre points the sampler uses (nect/sampling/ct_sampling.cu:
226-389).

Two families are provided:
  * ``ellipsoids`` -- Shepp-Logan-style random ellipsoids (attenuation 0..1);
  * ``packed_spheres`` -- non-overlapping random spheres (grains) inside a
    cylinder filled with pore fluid and surrounded by a wall.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch


@dataclass
class Phantom:
    centres: torch.Tensor   # [K, 3] x, y, z mm (float64)
    axes: torch.Tensor      # [K, 3] semi-axes mm
    rots: torch.Tensor      # [K, 3, 3] columns = ellipsoid axes in the volume frame
    mus: torch.Tensor       # [K] attenuation per mm (additive)

    def to(self, device) -> "Phantom":
        return Phantom(self.centres.to(device), self.axes.to(device), self.rots.to(device), self.mus.to(device))


def _rotation(a: float, b: float, c: float) -> np.ndarray:
    ca, sa, cb, sb, cc, sc = math.cos(a), math.sin(a), math.cos(b), math.sin(b), math.cos(c), math.sin(c)
    rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return rz @ ry @ rx


def ellipsoids(s_voxel_mm, n: int = 10, seed: int = 0) -> Phantom:
    """n random ellipsoids: centres within 0.8 of the field of view, semi-axes 0.05-0.4 of it."""
    rng = np.random.default_rng(seed)
    fov = np.array([s_voxel_mm[2], s_voxel_mm[1], s_voxel_mm[0]], np.float64)  # x, y, z
    centres = (rng.random((n, 3)) - 0.5) * 0.8 * fov * 0.5
    axes = (0.05 + 0.35 * rng.random((n, 3))) * fov * 0.5
    rots = np.stack([_rotation(*rng.random(3) * math.pi) for _ in range(n)])
    mus = rng.random(n)
    return Phantom(*(torch.from_numpy(a) for a in (centres, axes, rots, mus)))


def packed_spheres(n_voxel: int, d_voxel_mm: float, seed: int = 1, r_min: float = 4, r_max: float = 12,
                   target_porosity: float = 0.35, grain: float = 1.0, pore: float = 0.05, wall: float = 0.2,
                   max_spheres: int = 20000) -> Phantom:
    """Random sequential addition of non-overlapping spheres (radii in voxels) inside a cylinder.

    The cylinder (axis z, radius 0.45 of the volume width) is modelled as two
    concentric cylinders of large z-extent approximated by very elongated
    ellipsoids: the wall ring (outer - inner) and the pore fluid (inner), with
    grains adding (grain - pore).  Random sequential addition jams well before
    a porosity of 0.35; the packing stops at the target or when no sphere fits
    after many tries.
    """
    rng = np.random.default_rng(seed)
    L = n_voxel * d_voxel_mm
    R_in = 0.42 * L
    R_out = 0.45 * L
    half_h = 0.48 * L
    vol = math.pi * R_in * R_in * 2 * half_h
    solid = 0.0
    cs = np.zeros((0, 3))
    rs = np.zeros(0)
    rounds = 0
    while solid / vol < 1 - target_porosity and cs.shape[0] < max_spheres and rounds < 200:
        rounds += 1
        m = 4096
        r = (r_min + (r_max - r_min) * rng.random(m)) * d_voxel_mm
        rad = (R_in - r) * np.sqrt(rng.random(m))
        th = 2 * math.pi * rng.random(m)
        c = np.stack([rad * np.cos(th), rad * np.sin(th), (2 * rng.random(m) - 1) * (half_h - r)], 1)
        if cs.shape[0]:
            d2 = ((c[:, None, :] - cs[None, :, :]) ** 2).sum(-1)
            free = np.all(d2 >= (r[:, None] + rs[None, :]) ** 2, axis=1)
            c, r = c[free], r[free]
        acc_c, acc_r = [], []
        for i in range(c.shape[0]):
            if acc_c:
                ac = np.asarray(acc_c)
                if np.any(((ac - c[i]) ** 2).sum(1) < (np.asarray(acc_r) + r[i]) ** 2):
                    continue
            acc_c.append(c[i])
            acc_r.append(r[i])
            solid += 4 / 3 * math.pi * r[i] ** 3
            if solid / vol >= 1 - target_porosity or cs.shape[0] + len(acc_c) >= max_spheres:
                break
        if acc_c:
            cs = np.concatenate([cs, np.asarray(acc_c)])
            rs = np.concatenate([rs, np.asarray(acc_r)])
    cs = list(cs)
    rs = list(rs)
    K = len(cs)
    big = 1e4
    centres = [np.zeros(3), np.zeros(3), np.zeros(3)] + cs
    axes = [np.array([R_out, R_out, big]), np.array([R_in, R_in, big]), np.array([R_in, R_in, big])] + [
        np.array([r, r, r]) for r in rs]
    mus = [wall, -wall, pore] + [grain - pore] * K
    rots = np.tile(np.eye(3), (K + 3, 1, 1))
    ph = Phantom(torch.from_numpy(np.asarray(centres)), torch.from_numpy(np.asarray(axes)), torch.from_numpy(rots),
                 torch.from_numpy(np.asarray(mus, np.float64)))
    ph.z_clip = half_h  # cylinders end at |z| = half_h
    ph.porosity = 1 - solid / vol
    return ph


def ray_endpoints(mode: str, n_det, d_det, dso: float, dsd: float, angle: float, pixels: torch.Tensor,
                  reach: float = 0.0):
    """(P1, P2) [n, 3] float64 of detector pixels (row*cols + col), pixel centres, x/y/z mm."""
    rows, cols = n_det
    pixels = pixels.to(torch.int64)
    col = (pixels % cols).double()
    row = torch.div(pixels, cols, rounding_mode="floor").double()
    u = (col - (cols - 1) / 2) * d_det[1]
    v = (row - (rows - 1) / 2) * d_det[0]
    n = pixels.numel()
    dev = pixels.device
    if mode == "cone":
        p1 = torch.tensor([-dso, 0.0, 0.0], dtype=torch.float64, device=dev).expand(n, 3)
        p2 = torch.stack([torch.full((n,), dsd - dso, dtype=torch.float64, device=dev), -u, v], 1)
    else:
        p2 = torch.stack([torch.full((n,), reach, dtype=torch.float64, device=dev), -u, v], 1)
        p1 = torch.stack([torch.full((n,), -reach, dtype=torch.float64, device=dev), -u, v], 1)
    c, s = math.cos(angle), math.sin(angle)
    rz = torch.tensor([[c, -s, 0], [s, c, 0], [0, 0, 1]], dtype=torch.float64, device=dev)
    return p1 @ rz.T, p2 @ rz.T


def line_integrals(ph: Phantom, P1: torch.Tensor, P2: torch.Tensor, chunk: int = 256) -> torch.Tensor:
    """Exact line integrals [n] (float64) of the phantom along segments P1 -> P2."""
    d = P2 - P1
    seg = d.norm(dim=1)
    out = torch.zeros(P1.shape[0], dtype=torch.float64, device=P1.device)
    zc = getattr(ph, "z_clip", None)
    K = ph.centres.shape[0]
    for k0 in range(0, K, chunk):
        c = ph.centres[k0:k0 + chunk]
        a = ph.axes[k0:k0 + chunk]
        R = ph.rots[k0:k0 + chunk]
        mu = ph.mus[k0:k0 + chunk]
        q = torch.einsum("nd,kde->nke", P1, R) - torch.einsum("kd,kde->ke", c, R)[None]
        q = q / a[None]
        e = torch.einsum("nd,kde->nke", d, R) / a[None]
        A = (e * e).sum(-1)
        B = 2 * (q * e).sum(-1)
        C = (q * q).sum(-1) - 1
        disc = B * B - 4 * A * C
        ok = disc > 0
        sq = torch.sqrt(torch.where(ok, disc, torch.zeros_like(disc)))
        s0 = ((-B - sq) / (2 * A)).clamp(0, 1)
        s1 = ((-B + sq) / (2 * A)).clamp(0, 1)
        if zc is not None and k0 == 0:
            # the first three (cylinder) entries end at |z| = z_clip
            z0 = P1[:, 2:3]
            dz = d[:, 2:3]
            safe = torch.where(dz.abs() > 1e-12, dz, torch.full_like(dz, 1e-12))
            ta = ((-zc - z0) / safe)
            tb = ((zc - z0) / safe)
            lo = torch.minimum(ta, tb).clamp(0, 1)
            hi = torch.maximum(ta, tb).clamp(0, 1)
            inside = (z0.abs() <= zc) & (dz.abs() <= 1e-12)
            lo = torch.where(dz.abs() > 1e-12, lo, torch.zeros_like(lo))
            hi = torch.where(dz.abs() > 1e-12, hi, inside.double())
            ncyl = min(3, s0.shape[1])
            s0 = torch.cat([torch.maximum(s0[:, :ncyl], lo), s0[:, ncyl:]], 1)
            s1 = torch.cat([torch.minimum(s1[:, :ncyl], hi), s1[:, ncyl:]], 1)
        length = torch.where(ok, (s1 - s0).clamp_min(0), torch.zeros_like(s0))
        out += (length * mu[None]).sum(1) * seg
    return out


def project(ph: Phantom, mode: str, n_det, d_det, dso: float, dsd: float, angles, device="cuda",
            reach: float = 0.0) -> torch.Tensor:
    """Analytic projections [n_views, rows, cols] (float32) of the phantom."""
    ph = ph.to(device)
    if hasattr(ph, "z_clip") is False:
        pass
    rows, cols = n_det
    pix = torch.arange(rows * cols, device=device)
    out = []
    for ang in angles:
        P1, P2 = ray_endpoints(mode, n_det, d_det, dso, dsd, float(ang), pix, reach)
        out.append(line_integrals(ph, P1, P2).float().view(rows, cols))
    return torch.stack(out)


def voxelize(ph: Phantom, n_voxel, d_voxel, device="cuda") -> torch.Tensor:
    """Attenuation [z, y, x] at voxel centres (float32)."""
    ph = ph.to(device)
    nz, ny, nx = n_voxel
    dz, dy, dx = d_voxel
    zs = (torch.arange(nz, device=device, dtype=torch.float64) - (nz - 1) / 2) * dz
    ys = (torch.arange(ny, device=device, dtype=torch.float64) - (ny - 1) / 2) * dy
    xs = (torch.arange(nx, device=device, dtype=torch.float64) - (nx - 1) / 2) * dx
    vol = torch.zeros((nz, ny, nx), dtype=torch.float32, device=device)
    Y, X = torch.meshgrid(ys, xs, indexing="ij")
    zc = getattr(ph, "z_clip", None)
    for iz in range(nz):
        p = torch.stack([X.flatten(), Y.flatten(), torch.full_like(X.flatten(), zs[iz])], 1)
        acc = torch.zeros(p.shape[0], dtype=torch.float64, device=device)
        q = torch.einsum("nd,kde->nke", p, ph.rots) - torch.einsum("kd,kde->ke", ph.centres, ph.rots)[None]
        inside = ((q / ph.axes[None]) ** 2).sum(-1) <= 1
        if zc is not None:
            inside[:, :3] &= (p[:, 2:3].abs() <= zc)
        acc = (inside.double() * ph.mus[None]).sum(1)
        vol[iz] = acc.float().view(ny, nx)
    return vol


def psnr(estimate: torch.Tensor, truth: torch.Tensor) -> float:
    """PSNR (dB) with the truth's dynamic range as the peak (the evaluator's PSNR, nect/src/evaluation/evaluator.py:17-19)."""
    mse = torch.mean((estimate.double() - truth.double()) ** 2).item()
    rng = (truth.max() - truth.min()).item()
    return 10 * math.log10(rng * rng / max(mse, 1e-30))
