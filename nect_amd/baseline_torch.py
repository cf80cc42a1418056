"""The plain FP32 GPU path: the same training step in stock PyTorch eager ops.

This is the denominator the fused HIP path is measured against, not a
fallback: nothing in the package routes here implicitly, and it is only run
when asked for by name (``bench.py`` and the PSNR comparison).  It follows the
reference batch body (nect/trainers/base_trainer.py:517-653) with
tiny-cuda-nn replaced by plain torch modules:

  * sampler: the geometry of nect/sampling/ct_sampling.cu:117-576 vectorised
    over rays in float64 torch ops, points [B, S, 3] float32 zero-padded;
  * mask + compaction of all-zero points (base_trainer.py:523-528);
  * hash-grid encoding with index/weight rules identical to the HIP kernels,
    gathered with torch indexing, float32 table;
  * bias-free MLP as F.linear layers in float32;
  * line integral, L1/L2 loss, autograd backward, torch.optim.Adam.
"""
from __future__ import annotations

import math

import torch
import torch.nn.functional as F

from .grid import GridDesc
from .mlp import MlpDesc

PRIMES = (1, 2654435761, 805459861, 3674653429)
MASK32 = 0xFFFFFFFF


def _f32(x):
    return torch.as_tensor(x, dtype=torch.float32)


def sample_points(g, angle: float, ray_index: torch.Tensor, starting_ray_index: int, num_rays: int, S: int,
                  step_mm: float, uniform: bool, jitter: torch.Tensor, det_offset: torch.Tensor):
    """Voxel-mode sampler (cone or parallel) in torch: (points [B,S,3] f32, distances [B] f32)."""
    dev = ray_index.device
    f32, f64 = torch.float32, torch.float64
    rows, cols = g.nDetector
    B = num_rays
    t = torch.arange(B, device=dev) + starting_ray_index
    in_idx = t < ray_index.numel()
    ray = torch.where(in_idx, ray_index[t.clamp(max=ray_index.numel() - 1)], torch.zeros_like(t)).to(torch.int32)
    valid = in_idx & (ray < rows * cols)
    ray = torch.where(valid, ray, torch.zeros_like(ray))
    col = (ray % cols).to(f32)
    row = torch.div(ray, cols, rounding_mode="floor").to(f32)
    half_c = _f32(float(cols - 1)) / 2
    half_r = _f32(float(rows - 1)) / 2
    u = ((col - half_c.to(dev)) + det_offset[:, 0]) * _f32(g.dDetector[1]).to(dev) + _f32(g.offDetector[1]).to(dev)
    v = ((row - half_r.to(dev)) + det_offset[:, 1]) * _f32(g.dDetector[0]).to(dev) + _f32(g.offDetector[0]).to(dev)
    p2x = torch.zeros(B, dtype=f64, device=dev)
    p2y = -u.to(f64)
    p2z = v.to(f64)
    sv = [float(torch.tensor(x, dtype=f32)) for x in g.sVoxel]  # z, y, x as float32 values
    dv = [float(torch.tensor(x, dtype=f32)) for x in g.dVoxel]
    cone = hasattr(g, "distance_source_to_origin")
    if cone:
        p2x = p2x + float(torch.tensor(g.distance_origin_to_detector, dtype=f32))
        p1x = torch.full((B,), -float(torch.tensor(g.distance_source_to_origin, dtype=f32)), dtype=f64, device=dev)
        p1y = torch.zeros(B, dtype=f64, device=dev)
        p1z = torch.zeros(B, dtype=f64, device=dev)
    else:
        po = torch.sqrt(_f32(sv[2]) * _f32(sv[2]) + _f32(sv[1]) * _f32(sv[1]))
        p2x = p2x + float(po)
        p1x, p1y, p1z = -p2x, p2y.clone(), p2z.clone()
    c = float(torch.tensor(math.cos(float(torch.tensor(angle, dtype=f32))), dtype=f32))
    s = float(torch.tensor(math.sin(float(torch.tensor(angle, dtype=f32))), dtype=f32))
    cor = float(torch.tensor(g.COR, dtype=f32))
    p1y, p2y = p1y - cor, p2y - cor
    p1x, p1y = c * p1x - s * p1y, s * p1x + c * p1y
    p2x, p2y = c * p2x - s * p2y, s * p2x + c * p2y
    hx = float((_f32(sv[2]) - _f32(dv[2]) / 2) / 2)
    hy = float((_f32(sv[1]) - _f32(dv[1]) / 2) / 2)
    hz = float((_f32(sv[0]) - _f32(dv[0]) / 2) / 2)
    dx, dy, dz = p2x - p1x, p2y - p1y, p2z - p1z

    def slab(a, d, h):
        t0 = (-h - a) / d
        t1 = (h - a) / d
        return torch.minimum(t0, t1), torch.maximum(t0, t1)

    ax0, ax1 = slab(p1x, dx, hx)
    ay0, ay1 = slab(p1y, dy, hy)
    az0, az1 = slab(p1z, dz, hz)
    t0 = torch.maximum(torch.maximum(ax0, ay0), az0)
    t1 = torch.minimum(torch.minimum(ax1, ay1), az1)
    hit = valid & (t0 <= t1) & (t0 >= 0) & (t0 <= 1) & (t1 >= 0) & (t1 <= 1)
    sx, sy, sz = p1x + t0 * dx, p1y + t0 * dy, p1z + t0 * dz
    ex, ey, ez = p1x + t1 * dx, p1y + t1 * dy, p1z + t1 * dz
    stx, sty, stz = ex - sx, ey - sy, ez - sz
    rd = torch.sqrt(stx * stx + sty * sty + stz * stz).to(f32).to(f64)
    if uniform:
        dist = torch.full((B,), float(torch.tensor(step_mm, dtype=f32)), dtype=f32, device=dev)
        dmm = float(torch.tensor(step_mm, dtype=f32))
        stx, sty, stz = (stx / rd) * dmm / sv[2], (sty / rd) * dmm / sv[1], (stz / rd) * dmm / sv[0]
    else:
        dist = rd.to(f32) / S
        stx, sty, stz = (stx / S) / sv[2], (sty / S) / sv[1], (stz / S) / sv[0]
    sx, ex, sy, ey, sz, ez = sx / sv[2], ex / sv[2], sy / sv[1], ey / sv[1], sz / sv[0], ez / sv[0]
    j = jitter.to(f64)
    sx, sy, sz = (sx + 0.5) + j * stx, (sy + 0.5) + j * sty, (sz + 0.5) + j * stz
    ex, ey, ez = (ex + 0.5) + stx, (ey + 0.5) + sty, (ez + 0.5) + stz
    cl = lambda q: torch.nan_to_num(q, nan=0.0).clamp(0.0, 1.0)
    sx, sy, sz, ex, ey, ez = cl(sx), cl(sy), cl(sz), cl(ex), cl(ey), cl(ez)
    i = torch.arange(S, dtype=f64, device=dev)[None, :]
    px = (sx[:, None] + i * stx[:, None]).to(f32)
    py = (sy[:, None] + i * sty[:, None]).to(f32)
    pz = (sz[:, None] + i * stz[:, None]).to(f32)
    eps = 1e-5

    def brk(p, st, e):
        up = (st[:, None] > 0) & ((p - eps).to(f64) > e[:, None])
        dn = (st[:, None] < 0) & ((p + eps).to(f64) < e[:, None])
        return up | dn

    stop = brk(px, stx, ex) | brk(py, sty, ey) | brk(pz, stz, ez)
    written = (torch.cumsum(stop.to(torch.int32), 1) == 0) & hit[:, None]
    pts = torch.stack([pz, py, px], -1).clamp(0, 1)
    pts = torch.where(written[..., None], pts, torch.zeros((), dtype=f32, device=dev))
    dist = torch.where(hit, dist, torch.zeros((), dtype=f32, device=dev))
    return pts, dist


class TorchHashGrid(torch.nn.Module):
    """Hash-grid encoding in torch ops (float32 table), same rules as the HIP kernels."""

    def __init__(self, desc: GridDesc, table: torch.Tensor):
        super().__init__()
        self.desc = desc
        self.table = torch.nn.Parameter(table.view(-1, desc.n_features).clone())

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        d = self.desc
        D = d.n_dims
        out = []
        for l in range(d.n_levels):
            scale = torch.tensor(d.scale[l], dtype=torch.float32, device=x.device)
            p = (x * scale) + 0.5
            fl = torch.floor(p)
            frac = p - fl
            c = fl.to(torch.int64)
            res, size = d.resolution[l], d.level_size[l]
            acc = 0
            for k in range(1 << D):
                w = torch.ones(x.shape[0], dtype=torch.float32, device=x.device)
                cc = []
                for dd in range(D):
                    if (k >> dd) & 1:
                        w = w * frac[:, dd]
                        cc.append(c[:, dd] + 1)
                    else:
                        w = w * (1.0 - frac[:, dd])
                        cc.append(c[:, dd])
                stride, index, dense = 1, torch.zeros_like(cc[0]), True
                for dd in range(D):
                    if stride > size:
                        dense = False
                        break
                    index = (index + cc[dd] * stride) & MASK32
                    stride *= res
                if stride > size:
                    dense = False
                if not dense:
                    h = torch.zeros_like(cc[0])
                    for dd in range(D):
                        h = h ^ ((cc[dd] * PRIMES[dd]) & MASK32)
                    index = h
                index = index % size + d.offset[l]
                acc = acc + w[:, None] * self.table[index]
            out.append(acc)
        return torch.cat(out, 1)


class TorchMlp(torch.nn.Module):
    def __init__(self, desc: MlpDesc, weights: torch.Tensor):
        super().__init__()
        self.desc = desc
        self.weights = torch.nn.ParameterList([torch.nn.Parameter(w.clone()) for w in desc.split(weights)])

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        from .tcnn import torch_activation

        d = self.desc
        if x.shape[1] < d.n_input_padded:
            x = F.pad(x, (0, d.n_input_padded - x.shape[1]))
        h = x
        for w in self.weights[:-1]:
            h = torch_activation(d.activation, F.linear(h, w))
        return torch_activation(d.output_activation, F.linear(h, self.weights[-1])[:, : d.n_output])


class TorchTrainStep:
    """The reference batch body on stock torch ops in float32, for timing and PSNR comparisons."""

    def __init__(self, grid: GridDesc, mlp: MlpDesc, params: torch.Tensor, lr: float = 1e-2, loss: str = "L2",
                 betas=(0.9, 0.999), eps: float = 1e-8):
        n_mlp = mlp.n_params
        self.encoder = TorchHashGrid(grid, params[n_mlp:].float())
        self.mlp = TorchMlp(mlp, params[:n_mlp].float())
        self.opt = torch.optim.Adam(list(self.mlp.parameters()) + list(self.encoder.parameters()), lr=lr,
                                    betas=betas, eps=eps)
        self.loss = "l1" if "l1" in loss.casefold() else "l2"

    def model(self, x):
        return self.mlp(self.encoder(x))

    def step(self, geometry, angle, ray_index, starting_ray_index, num_rays, S, step_mm, uniform, jitter, det_offset,
             target, l_max, lr=None):
        if lr is not None:
            for g in self.opt.param_groups:
                g["lr"] = lr
        self.opt.zero_grad()
        points, dist = sample_points(geometry, angle, ray_index, starting_ray_index, num_rays, S, step_mm, uniform,
                                     jitter, det_offset)
        mask = torch.all(points.view(-1, 3) == 0, dim=-1)
        kept = points.view(-1, 3)[~mask]
        mu = self.model(kept)
        full = torch.zeros((num_rays * S, 1), dtype=torch.float32, device=points.device)
        full[~mask] = mu
        y_pred = full.view(num_rays, S).sum(1) * (dist / l_max)
        if self.loss == "l1":
            loss = F.l1_loss(y_pred, target)
        else:
            loss = F.mse_loss(y_pred, target)
        loss.backward()
        self.opt.step()
        return loss.detach()

    def flat_params(self) -> torch.Tensor:
        return torch.cat([w.detach().flatten() for w in self.mlp.weights] + [self.encoder.table.detach().flatten()])


class TorchExport:
    """Sequential float32 export queries on stock torch ops: one request at a time, in batches of up to
    5M voxels, as the reference's export loop evaluates them (nect/static_export.py:106-141); the
    coordinates are built on the device, the rescale is ((mu / l_max) * span) + lo, and the result is copied
    to the host as the reference's exports end (.cpu())."""

    def __init__(self, grid: GridDesc, mlp: MlpDesc, params: torch.Tensor, batch: int = 5_000_000):
        n_mlp = mlp.n_params
        self.grid = grid
        self.encoder = TorchHashGrid(grid, params[n_mlp:].float())
        self.mlp = TorchMlp(mlp, params[:n_mlp].float())
        self.batch = batch

    @torch.no_grad()
    def run(self, r) -> torch.Tensor:
        dev = self.encoder.table.device
        axes = [torch.linspace(r.a[k], r.b[k], steps=r.n[k], device=dev) for k in range(3)]
        out = torch.empty(r.voxels, dtype=torch.float32, device=dev)
        ny, nx = r.n[1], r.n[2]
        for v0 in range(0, r.voxels, self.batch):
            i = torch.arange(v0, min(v0 + self.batch, r.voxels), device=dev)
            cols = [axes[0][i // (ny * nx)], axes[1][(i // nx) % ny], axes[2][i % nx]]
            if self.grid.n_dims == 4:
                cols.append(torch.full_like(cols[0], r.t))
            y = self.mlp(self.encoder(torch.stack(cols, 1)))[:, 0]
            out[v0: v0 + y.shape[0]] = ((y / r.l_max) * r.span) + r.lo
        return out.view(r.n).cpu()

    def serve(self, requests) -> dict:
        """Replays requests at their arrival times and serves them one at a time, in arrival order."""
        import time

        from .serving import summarize

        t0 = time.perf_counter()
        done = []
        for r in sorted(requests, key=lambda q: q.arrival):
            wait = r.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            r.admitted = time.perf_counter() - t0
            self.run(r)
            torch.cuda.synchronize()
            r.done = time.perf_counter() - t0
            done.append(r)
        return summarize(done, time.perf_counter() - t0)
