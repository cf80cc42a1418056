"""Seeded synthetic phantoms and their analytic projections (data generation, not the hot path).

A phantom is a sum of ellipsoids (attenuation per mm, additive) in the volume
frame of the sampler: millimetres, volume centred on the origin, x along the
source-detector axis at angle 0, z along detector rows.  Projections are exact
line integrals computed in float64 with torch (on the GPU when the rays are
there): the chord of each ellipsoid follows from a quadratic in its unit-sphere
frame.  The ray of detector pixel (row, col) joins the same source and
detector-pixel-centre points the sampler uses (nect/sampling/ct_sampling.cu:
226-389).

Families:
  * ``ellipsoids`` -- Shepp-Logan-style random ellipsoids (attenuation 0..1);
  * ``packed_spheres`` / ``relaxed_packed_spheres`` -- random spheres (grains)
    inside a cylinder filled with pore fluid and surrounded by a wall (the
    relaxed packing reaches porosity 0.35; its grain phase is the union of the
    spheres, projected exactly by ``union_chords``);
  * ``brine_front`` -- the relaxed packing in time: brine below a rising front,
    soluble grains shrinking as it passes (each view at its own t);
  * ``ellipse_slab`` -- random elliptic cylinders of random z-extent in a slab
    (about 200 crossing each slice).
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
        out = Phantom(self.centres.to(device), self.axes.to(device), self.rots.to(device), self.mus.to(device))
        for k, v in vars(self).items():
            if k not in ("centres", "axes", "rots", "mus"):
                setattr(out, k, v.to(device) if isinstance(v, torch.Tensor) else v)
        return out


def ellipse_slab(n_xy: int = 512, n_z: int = 64, d_voxel_mm: float = 1.0, per_slice: int = 200,
                 seed: int = 4) -> Phantom:
    """A slab of n_xy x n_xy x n_z voxels holding random ellipses extruded over random z-extents (elliptic
    cylinders with axis z): semi-axes 4-40 voxels, any orientation in the plane, centres anywhere in the
    slab's square, extents of 8 to n_z slices; attenuation U(0, 1) each, additive where they overlap; enough of
    them that about ``per_slice`` cross each slice."""
    rng = np.random.default_rng(seed)
    half_xy, half_z = 0.5 * n_xy * d_voxel_mm, 0.5 * n_z * d_voxel_mm
    ext = rng.integers(8, n_z + 1, 1 << 14)
    n = int(round(per_slice * n_z / ext.mean()))
    ext = ext[:n]
    z0 = rng.integers(0, n_z - ext + 1)
    zr = np.stack([z0, z0 + ext], 1) * d_voxel_mm - half_z
    ab = (4 + 36 * rng.random((n, 2))) * d_voxel_mm
    xy = (rng.random((n, 2)) * 2 - 1) * half_xy
    centres = np.concatenate([xy, zr.mean(1, keepdims=True)], 1)
    axes = np.concatenate([ab, np.full((n, 1), 1e4)], 1)
    th = rng.random(n) * math.pi
    rots = np.zeros((n, 3, 3))
    rots[:, 0, 0], rots[:, 0, 1], rots[:, 1, 0], rots[:, 1, 1] = np.cos(th), -np.sin(th), np.sin(th), np.cos(th)
    rots[:, 2, 2] = 1.0
    ph = Phantom(*(torch.from_numpy(a) for a in (centres, axes, rots, rng.random(n))))
    ph.z_ranges = torch.from_numpy(zr.astype(np.float64))
    return ph


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


def relaxed_packed_spheres(n_voxel: int, d_voxel_mm: float, seed: int = 1, r_min: float = 4, r_max: float = 12,
                           target_porosity: float = 0.35, grain: float = 1.0, pore: float = 0.05, wall: float = 0.2,
                           iterations: int = 300, device="cpu") -> Phantom:
    """Grains (spheres, radii U(r_min, r_max) voxels) packed to about a target porosity in the cylinder of
    ``packed_spheres`` (radius 0.42 of the volume width, wall to 0.45, height 0.96 of it).

    Spheres whose volumes add up to (1 - target_porosity) of the cylinder are placed at random inside it and
    relaxed by soft-sphere repulsion (each overlapping pair pushed apart by half its overlap per iteration,
    spheres kept inside the cylinder) for ``iterations`` rounds; the grain phase is the union of the spheres
    (what overlap remains is fused grain, attenuation ``grain`` everywhere in it), so the porosity is that of
    the union -- 0.35 for the default target on 128^3-256^3 volumes (see ``union_porosity``).  Random
    sequential addition without overlap (``packed_spheres``) jams far below this density.  The phantom holds
    the wall and pore cylinders as ellipsoids and the grains as ``sphere_centres``/``sphere_radii``;
    ``project`` and ``voxelize`` take the union into account."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    f64 = torch.float64
    L = n_voxel * d_voxel_mm
    R_in, R_out, half_h = 0.42 * L, 0.45 * L, 0.48 * L
    vol = math.pi * R_in * R_in * 2 * half_h
    mean_r3 = ((r_max ** 4 - r_min ** 4) / (4 * (r_max - r_min))) * d_voxel_mm ** 3
    n = int(round((1 - target_porosity) * vol / (4 / 3 * math.pi * mean_r3)))
    r = (r_min + (r_max - r_min) * torch.rand(n, generator=g, dtype=f64)) * d_voxel_mm
    rad = (R_in - r) * torch.sqrt(torch.rand(n, generator=g, dtype=f64))
    th = 2 * math.pi * torch.rand(n, generator=g, dtype=f64)
    z = (2 * torch.rand(n, generator=g, dtype=f64) - 1) * (half_h - r)
    c = torch.stack([rad * torch.cos(th), rad * torch.sin(th), z], 1).to(device)
    r = r.to(device)
    eye = torch.eye(n, dtype=torch.bool, device=device)
    lim = (R_in - r).clamp_min(0)
    for _ in range(iterations):
        d = c[:, None, :] - c[None, :, :]
        dist = d.norm(dim=-1).masked_fill(eye, float("inf"))
        ov = (r[:, None] + r[None, :] - dist).clamp_min(0)
        c = c + ((ov / dist.clamp_min(1e-12) * 0.5)[..., None] * d).sum(1)
        rr = c[:, :2].norm(dim=1)
        scale = torch.where(rr > lim, lim / rr.clamp_min(1e-12), torch.ones_like(rr))
        c = torch.cat([c[:, :2] * scale[:, None], c[:, 2:3].clamp(-(half_h - r)[:, None], (half_h - r)[:, None])], 1)
    centres = torch.zeros((3, 3), dtype=f64)
    axes = torch.tensor([[R_out, R_out, 1e4], [R_in, R_in, 1e4], [R_in, R_in, 1e4]], dtype=f64)
    mus = torch.tensor([wall, -wall, pore], dtype=f64)
    ph = Phantom(centres, axes, torch.eye(3, dtype=f64).expand(3, 3, 3).contiguous(), mus)
    ph.z_clip = half_h
    ph.sphere_centres, ph.sphere_radii = c.cpu(), r.cpu()
    ph.grain_excess = grain - pore
    ph.cylinder = (R_in, R_out, half_h)
    return ph


def _below_interval(P1: torch.Tensor, d: torch.Tensor, z_below: float):
    """Per segment P1 + s*d (s in [0, 1]), the parameter interval [lo, hi] where z < z_below."""
    z0, dz = P1[:, 2], d[:, 2]
    flat = dz.abs() <= 1e-12
    cross = (z_below - z0) / torch.where(flat, torch.ones_like(dz), dz)
    lo = torch.where(dz < -1e-12, cross, torch.zeros_like(z0)).clamp(0, 1)
    hi = torch.where(dz > 1e-12, cross, torch.ones_like(z0)).clamp(0, 1)
    empty = flat & (z0 >= z_below)
    return lo, torch.where(empty, lo, hi)


def union_chords(centres: torch.Tensor, radii: torch.Tensor, P1: torch.Tensor, P2: torch.Tensor,
                 chunk: int = 1024, z_below: float | None = None) -> torch.Tensor:
    """Length (mm, float64) of the segment P1 -> P2 inside the union of spheres (within the half-space
    z < z_below when given): each sphere's chord interval, the intervals sorted by entry, and the parts beyond
    the running maximum of the exits summed."""
    d = P2 - P1
    seg = d.norm(dim=1)
    u = d / seg[:, None]
    C = centres.to(P1.device)
    r2 = (radii.to(P1.device) ** 2)[None]
    out = torch.zeros(P1.shape[0], dtype=torch.float64, device=P1.device)
    if C.shape[0] == 0:
        return out
    if z_below is not None:
        blo, bhi = _below_interval(P1, d, z_below)
        blo, bhi = blo * seg, bhi * seg
    for i in range(0, P1.shape[0], chunk):
        w = P1[i:i + chunk, None, :] - C[None]
        b = (w * u[i:i + chunk, None, :]).sum(-1)
        disc = b * b - ((w * w).sum(-1) - r2)
        hit = disc > 0
        sq = torch.sqrt(disc.clamp_min(0))
        L = seg[i:i + chunk, None]
        lo = torch.zeros_like(L) if z_below is None else blo[i:i + chunk, None]
        hi = L if z_below is None else bhi[i:i + chunk, None]
        t0 = torch.where(hit, torch.minimum(torch.maximum(-b - sq, lo), hi), hi)
        t1 = torch.where(hit, torch.minimum(torch.maximum(-b + sq, lo), hi), hi)
        t0s, idx = torch.sort(t0, dim=1)
        t1s = torch.gather(t1, 1, idx)
        run = torch.cummax(t1s, dim=1).values
        prev = torch.cat([torch.zeros_like(run[:, :1]), run[:, :-1]], 1)
        out[i:i + chunk] = (t1s - torch.maximum(t0s, prev)).clamp_min(0).sum(1)
    return out


def union_chords_by_row(centres: torch.Tensor, radii: torch.Tensor, P1: torch.Tensor, P2: torch.Tensor,
                        cols: int, rows_per_block: int = 4, z_below: float | None = None) -> torch.Tensor:
    """union_chords for the rays of whole detector rows (rays row-major, ``cols`` per row): the rays of a row
    lie in one plane, so only spheres within their radius of a block's row planes can be hit; each block of
    rows is measured against those alone (same result, a fraction of the work)."""
    n = P1.shape[0]
    out = torch.zeros(n, dtype=torch.float64, device=P1.device)
    C = centres.to(P1.device)
    R = radii.to(P1.device)
    step = cols * rows_per_block
    for i in range(0, n, step):
        j = min(n, i + step)
        near = torch.zeros(C.shape[0], dtype=torch.bool, device=P1.device)
        for k in range(i, j, cols):
            a, b, c = P1[k], P2[k], P2[min(k + cols, j) - 1]
            nrm = torch.cross(b - a, c - a, dim=0)
            if float(nrm.norm()) < 1e-12:       # a one-pixel row: its single ray decides nothing
                near[:] = True
                break
            near |= (((C - a) @ (nrm / nrm.norm())).abs() <= R + 1e-9)
        if bool(near.any()):
            out[i:j] = union_chords(C[near], R[near], P1[i:j], P2[i:j], chunk=step, z_below=z_below)
    return out


def _cylinder_chord(P1: torch.Tensor, P2: torch.Tensor, radius: float, z_lo: float, z_hi: float) -> torch.Tensor:
    """Length (mm) of P1 -> P2 inside the z-aligned cylinder x^2 + y^2 <= radius^2, z_lo <= z <= z_hi."""
    d = P2 - P1
    seg = d.norm(dim=1)
    a = d[:, 0] ** 2 + d[:, 1] ** 2
    b = 2 * (P1[:, 0] * d[:, 0] + P1[:, 1] * d[:, 1])
    c = P1[:, 0] ** 2 + P1[:, 1] ** 2 - radius * radius
    disc = b * b - 4 * a * c
    ok = (disc > 0) & (a > 1e-18)
    sq = torch.sqrt(disc.clamp_min(0))
    aa = torch.where(ok, a, torch.ones_like(a))
    s0 = ((-b - sq) / (2 * aa)).clamp(0, 1)
    s1 = ((-b + sq) / (2 * aa)).clamp(0, 1)
    if z_hi <= z_lo:
        return torch.zeros_like(seg)
    lo1, hi1 = _below_interval(P1, d, z_hi)
    neg = P1.clone()
    neg[:, 2] = -P1[:, 2]
    dneg = d.clone()
    dneg[:, 2] = -d[:, 2]
    lo2, hi2 = _below_interval(neg, dneg, -z_lo)     # z > z_lo  <=>  -z < -z_lo
    lo = torch.maximum(torch.maximum(s0, lo1), lo2)
    hi = torch.minimum(torch.minimum(s1, hi1), hi2)
    return torch.where(ok, (hi - lo).clamp_min(0), torch.zeros_like(seg)) * seg


def brine_front(ph: Phantom, n_voxel: int, d_voxel_mm: float, soluble_fraction: float = 0.2, brine: float = 0.3,
                seed: int = 2) -> Phantom:
    """The time-dependent sample of a ``relaxed_packed_spheres`` phantom: brine (attenuation ``brine``) replaces
    the pore fluid below a planar front that rises from the bottom of the volume (z = -n*d/2) at t = 0 to its top
    at t = 1; a seeded ``soluble_fraction`` of the grains shrink linearly to zero radius while the front passes
    them, radius r0 * clamp((z_c + r0 - h(t)) / (2 r0), 0, 1) at front height h(t).  ``at_time`` gives the
    phantom at one t; ``project`` takes each view's t."""
    out = scaled(ph, 1.0)
    g = torch.Generator().manual_seed(seed)
    out.soluble = torch.rand(ph.sphere_radii.numel(), generator=g, dtype=torch.float64) < soluble_fraction
    out.brine_excess = brine - float(ph.mus[2])
    half = 0.5 * n_voxel * d_voxel_mm
    out.front = (-half, half)
    return out


def front_height(ph: Phantom, t: float) -> float:
    z0, z1 = ph.front
    return z0 + float(t) * (z1 - z0)


def at_time(ph: Phantom, t: float) -> Phantom:
    """The brine-front phantom's grains at time t (the soluble ones shrunk), with ``brine_level`` = h(t)."""
    out = scaled(ph, 1.0)
    h = front_height(ph, t)
    r0, zc = ph.sphere_radii, ph.sphere_centres[:, 2]
    shrunk = r0 * ((zc + r0 - h) / (2 * r0)).clamp(0, 1)
    out.sphere_radii = torch.where(ph.soluble, shrunk, r0)
    out.brine_level = h
    return out


def union_porosity(ph: Phantom, n_voxel: int, d_voxel_mm: float, device="cpu") -> float:
    """Fraction of the pore cylinder's voxels (centres of an n^3 grid) outside every grain."""
    vol = voxelize_spheres(ph, n_voxel, d_voxel_mm, device=device, levels=(1.0, 0.0, 0.0))
    n = n_voxel
    coords = (torch.arange(n, device=device, dtype=torch.float64) - (n - 1) / 2) * d_voxel_mm
    Z, Y, X = torch.meshgrid(coords, coords, coords, indexing="ij")
    R_in, _, half_h = ph.cylinder
    incyl = (X * X + Y * Y <= R_in * R_in) & (Z.abs() <= half_h)
    return 1.0 - float((vol[incyl] > 0).sum()) / float(incyl.sum())


def scaled(ph: Phantom, factor: float) -> Phantom:
    """The phantom with every attenuation multiplied by ``factor``."""
    out = Phantom(ph.centres, ph.axes, ph.rots, ph.mus * factor)
    for k, v in vars(ph).items():
        if k not in ("centres", "axes", "rots", "mus"):
            setattr(out, k, v)
    if hasattr(ph, "grain_excess"):
        out.grain_excess = ph.grain_excess * factor
    if hasattr(ph, "brine_excess"):
        out.brine_excess = ph.brine_excess * factor
    return out


def poisson_log(proj: torch.Tensor, i0: float, generator: torch.Generator | None = None) -> torch.Tensor:
    """Noisy line integrals: counts ~ Poisson(i0 * exp(-p)) per pixel, then -log(counts / i0); a pixel that
    counts no photon is given one count."""
    counts = torch.poisson(i0 * torch.exp(-proj.double()), generator=generator)
    return (-torch.log(counts.clamp_min(1.0) / i0)).float()


def voxelize_spheres(ph: Phantom, n_voxel: int, d_voxel_mm: float, device="cuda", levels=None) -> torch.Tensor:
    """Attenuation [z, y, x] at the voxel centres of an n^3 grid for a ``relaxed_packed_spheres`` phantom: the
    wall and pore cylinders, plus the grain excess over the union of the spheres (each sphere rasterised in
    its bounding box), plus the brine below the front for a brine-front phantom at one time (``at_time``).
    ``levels`` = (grain excess, pore, wall) overrides the phantom's attenuations (and leaves out the brine)."""
    n = n_voxel
    coords = (torch.arange(n, device=device, dtype=torch.float64) - (n - 1) / 2) * d_voxel_mm
    R_in, R_out, half_h = ph.cylinder
    if levels is None:
        excess, pore, wall = ph.grain_excess, float(ph.mus[2]), float(ph.mus[0])
    else:
        excess, pore, wall = levels
    vol = torch.zeros((n, n, n), dtype=torch.float32, device=device)
    grains = torch.zeros((n, n, n), dtype=torch.bool, device=device)
    Y, X = torch.meshgrid(coords, coords, indexing="ij")
    rho2 = X * X + Y * Y
    plane = ((rho2 <= R_out * R_out).float() * wall - (rho2 <= R_in * R_in).float() * wall
             + (rho2 <= R_in * R_in).float() * pore)
    inz = (coords.abs() <= half_h).float()
    vol += inz[:, None, None] * plane[None].float()
    for (cx, cy, cz), r in zip(ph.sphere_centres.tolist(), ph.sphere_radii.tolist()):
        lo = [max(0, int(math.floor((v - r) / d_voxel_mm + (n - 1) / 2))) for v in (cz, cy, cx)]
        hi = [min(n, int(math.ceil((v + r) / d_voxel_mm + (n - 1) / 2)) + 1) for v in (cz, cy, cx)]
        zz = coords[lo[0]:hi[0]][:, None, None]
        yy = coords[lo[1]:hi[1]][None, :, None]
        xx = coords[lo[2]:hi[2]][None, None, :]
        grains[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] |= (zz - cz) ** 2 + (yy - cy) ** 2 + (xx - cx) ** 2 <= r * r
    vol += grains.float() * excess
    h = getattr(ph, "brine_level", None)
    if h is not None and levels is None:
        # brine below the front: pore cylinder, z < h, outside the grains
        pore_disc = (rho2 <= R_in * R_in)
        zmask = (coords < h) & (coords.abs() <= half_h)
        vol += (zmask[:, None, None] & pore_disc[None] & ~grains).float() * ph.brine_excess
    return vol


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
        zr = getattr(ph, "z_ranges", None)
        if zr is not None:
            # entries that end at their own z range [z_lo, z_hi]
            z0, dz = P1[:, 2:3], d[:, 2:3]
            zlo, zhi = zr[k0:k0 + chunk, 0][None], zr[k0:k0 + chunk, 1][None]
            flat = dz.abs() <= 1e-12
            safe = torch.where(flat, torch.ones_like(dz), dz)
            ta, tb = (zlo - z0) / safe, (zhi - z0) / safe
            lo = torch.where(flat, torch.zeros_like(ta), torch.minimum(ta, tb)).clamp(0, 1)
            within = ((z0 >= zlo) & (z0 <= zhi)).double()
            hi = torch.where(flat, within, torch.maximum(ta, tb).clamp(0, 1))
            s0 = torch.maximum(s0, lo)
            s1 = torch.minimum(s1, hi)
        length = torch.where(ok, (s1 - s0).clamp_min(0), torch.zeros_like(s0))
        out += (length * mu[None]).sum(1) * seg
    return out


def project(ph: Phantom, mode: str, n_det, d_det, dso: float, dsd: float, angles, device="cuda",
            reach: float = 0.0, times=None) -> torch.Tensor:
    """Analytic projections [n_views, rows, cols] (float32) of the phantom; a ``brine_front`` phantom is seen
    at each view's time ``times[v]``."""
    ph = ph.to(device)
    rows, cols = n_det
    pix = torch.arange(rows * cols, device=device)
    out = []
    spheres = getattr(ph, "sphere_radii", None)
    for v, ang in enumerate(angles):
        P1, P2 = ray_endpoints(mode, n_det, d_det, dso, dsd, float(ang), pix, reach)
        p = line_integrals(ph, P1, P2)
        if spheres is not None:
            q = at_time(ph, times[v]) if times is not None else ph
            p = p + union_chords_by_row(q.sphere_centres, q.sphere_radii, P1, P2, cols,
                                        rows_per_block=16) * ph.grain_excess
            if times is not None:
                # brine below the front: the pore cylinder there, less the grains there
                R_in, _, half_h = ph.cylinder
                h = q.brine_level
                below = _cylinder_chord(P1, P2, R_in, -half_h, min(h, half_h))
                grains = union_chords_by_row(q.sphere_centres, q.sphere_radii, P1, P2, cols, rows_per_block=16,
                                             z_below=h)
                p = p + (below - grains) * ph.brine_excess
        out.append(p.float().view(rows, cols))
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
        zr = getattr(ph, "z_ranges", None)
        if zr is not None:
            inside &= ((zs[iz] >= zr[:, 0]) & (zs[iz] <= zr[:, 1]))[None]
        acc = (inside.double() * ph.mus[None]).sum(1)
        vol[iz] = acc.float().view(ny, nx)
    return vol


def psnr(estimate: torch.Tensor, truth: torch.Tensor) -> float:
    """PSNR (dB) with the truth's dynamic range as the peak (the evaluator's PSNR, nect/src/evaluation/evaluator.py:17-19)."""
    mse = torch.mean((estimate.double() - truth.double()) ** 2).item()
    rng = (truth.max() - truth.min()).item()
    return 10 * math.log10(rng * rng / max(mse, 1e-30))
