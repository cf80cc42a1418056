"""Shared test helpers: one geometry description -> product struct and oracle fields."""
from __future__ import annotations

import numpy as np

from oracle.geometry import OracleGeometry


def product_struct(fields: dict):
    """The nect_amd ct_sampling geometry object for oracle struct fields."""
    from nect_amd.sampling import ct_sampling as cs

    common = dict(nDetector=fields["n_det"], dDetector=fields["d_det"], offDetector=fields["off_det"],
                  rotDetector=fields["rot_det"], COR=fields["cor"])
    k = fields["kind"]
    if k == 0:
        return cs.ConeGeometryVoxel(fields["dso"], fields["dod"], sVoxel=fields["s_voxel"], dVoxel=fields["d_voxel"],
                                    **common)
    if k == 1:
        return cs.ParallelGeometryVoxel(sVoxel=fields["s_voxel"], dVoxel=fields["d_voxel"], **common)
    if k == 2:
        return cs.ConeGeometryCylindrical(fields["dso"], fields["dod"], object_radius=fields["radius"],
                                          object_height=fields["height"], remove_factor_top=fields["remove_top"],
                                          remove_factor_bottom=fields["remove_bottom"], **common)
    return cs.ParallelGeometryCylindrical(object_radius=fields["radius"], object_height=fields["height"],
                                          remove_factor_top=fields["remove_top"],
                                          remove_factor_bottom=fields["remove_bottom"], **common)


def cfg1_geometry() -> OracleGeometry:
    """64^3 cone: detector 64x64 @ 1 mm, SOD 500, SDD 1000, 0.5 mm voxels."""
    return OracleGeometry(mode="cone", nDetector=(64, 64), dDetector=(1.0, 1.0), DSD=1000.0, DSO=500.0,
                          nVoxel=(64, 64, 64), dVoxel=(0.5, 0.5, 0.5))


def cfg2_geometry() -> OracleGeometry:
    """256^3 cone: detector 256x256 @ 1 mm, SOD 400, SDD 800, 0.5 mm voxels."""
    return OracleGeometry(mode="cone", nDetector=(256, 256), dDetector=(1.0, 1.0), DSD=800.0, DSO=400.0,
                          nVoxel=(256, 256, 256), dVoxel=(0.5, 0.5, 0.5))


def cfg4_geometry() -> OracleGeometry:
    """Parallel slab 512x512x64: detector 64 rows x 512 cols @ 1 mm, 1 mm voxels."""
    return OracleGeometry(mode="parallel", nDetector=(64, 512), dDetector=(1.0, 1.0), nVoxel=(64, 512, 512),
                          dVoxel=(1.0, 1.0, 1.0))


def seeded_batch(n_pixels: int, B: int, seed: int, det_offset_scale: float = 0.0):
    rng = np.random.default_rng(seed)
    idx = rng.permutation(n_pixels).astype(np.int64)
    jit = (rng.random(B, dtype=np.float32) - np.float32(0.5)).astype(np.float32)
    off = ((rng.random((B, 2), dtype=np.float32) - np.float32(0.5)) * np.float32(det_offset_scale)).astype(np.float32)
    return idx, jit, off


class FdCase:
    """Parallel-beam slab case for finite-difference checks of the training-step gradient: hash grid
    16x2 T=2^20 (growth 1.5, base 16), MLP 2x64 ReLU/softplus, L2 loss, a few rays of one view."""

    def __init__(self, seed: int = 0, B: int = 8, S: int = 192):
        import torch

        from oracle import hashgrid as hg
        from oracle.sampler import oracle_sample

        self.geometry = cfg4_geometry()
        self.fields = self.geometry.c_struct_fields()
        self.B, self.S = B, S
        self.l_max = self.geometry.max_distance_traveled
        self.step_mm = self.l_max / S
        self.angle = 0.7
        rng = np.random.default_rng(seed)
        # rays through the middle rows of the detector, so that every ray crosses the slab
        self.idx = (rng.integers(16, 48, B) * 512 + rng.integers(64, 448, B)).astype(np.int64)
        self.jit = (rng.random(B, dtype=np.float32) - np.float32(0.5)).astype(np.float32)
        self.off = np.zeros((B, 2), np.float32)
        self.spec = hg.GridSpec(3, 16, 2, 20, 16, 1.5)
        self.cfg = {"n_params": 64 * 32 + 64 * 64 + 16 * 64, "n_input_padded": 32, "n_hidden_layers": 2,
                    "n_output_padded": 16, "activation": "ReLU", "output_activation": "Softplus"}
        n_mlp = self.cfg["n_params"]
        w = np.zeros(n_mlp)
        w[: 64 * 32] = rng.normal(0, 1 / np.sqrt(32), 64 * 32)
        w[64 * 32: 64 * 32 + 64 * 64] = rng.normal(0, 1 / np.sqrt(64), 64 * 64)
        w[64 * 32 + 64 * 64: 64 * 32 + 64 * 64 + 64] = rng.normal(0, 1 / np.sqrt(64), 64)
        table = rng.uniform(-0.5, 0.5, self.spec.n_params)
        # float32-representable parameters, so the FP32 GPU path sees exactly these values
        self.params = np.concatenate([w, table]).astype(np.float32).astype(np.float64)
        self.target = (rng.random(B) * 2.0).astype(np.float32)
        pts, dist, cnt = oracle_sample(self.fields, self.angle, self.idx, B, 0, S, self.step_mm, True, self.jit,
                                       self.off)
        flat = pts.reshape(-1, 3)
        self.keep = ~np.all(flat == 0, axis=1)
        self.x = flat[self.keep]
        self.dist = dist
        self.counts = cnt
        self.corners = hg.corners(self.spec, self.x)
        self.p = torch.from_numpy(self.params.copy())

    def loss(self) -> float:
        import torch

        from oracle.train import integrate, model_forward

        with torch.no_grad():
            mu = model_forward(self.spec, self.cfg, self.p, self.x, corners=self.corners)
            full = torch.zeros(self.B * self.S, dtype=torch.float64)
            full[torch.from_numpy(self.keep)] = mu
            y = integrate(full.view(self.B, self.S), self.dist, self.l_max)
            return float(torch.mean((y - torch.from_numpy(self.target.astype(np.float64))) ** 2))

    def touched_entries(self) -> np.ndarray:
        """Parameter indices of the hash-table entries the samples touch."""
        idx, _ = self.corners
        rows = [self.spec.offset[l] + idx[:, l, :].astype(np.int64).ravel() for l in range(self.spec.n_levels)]
        ent = np.unique(np.concatenate(rows))
        F = self.spec.n_features
        return (self.cfg["n_params"] + (ent[:, None] * F + np.arange(F)[None]).ravel()).astype(np.int64)

    def _mlp_losses(self, weights) -> "torch.Tensor":
        """Losses [K] for K MLP weight vectors [K, n_mlp] at the current table (batched float64 forward)."""
        import torch

        from oracle import mlp as om
        from oracle.train import integrate, model_forward

        n_mlp = self.cfg["n_params"]
        with torch.no_grad():
            # features of the current table: the oracle forward with an identity-like probe is not available,
            # so gather them from the oracle's own corner data
            idx, wts = self.corners
            table = self.p[n_mlp:].view(-1, self.spec.n_features)
            feats = []
            for l in range(self.spec.n_levels):
                rows = torch.from_numpy(self.spec.offset[l] + idx[:, l, :].astype(np.int64))
                vals = table[rows.reshape(-1)].view(rows.shape[0], rows.shape[1], -1)
                feats.append((torch.from_numpy(wts[:, l, :].astype(np.float64))[:, :, None] * vals).sum(1))
            f = torch.cat(feats, 1)
            out = []
            keep = torch.from_numpy(self.keep)
            tgt = torch.from_numpy(self.target.astype(np.float64))
            for w in weights:
                W = om.split(w, 32, 64, 2, 16)
                mu = om.forward(W, f, "ReLU", "Softplus", 1)[:, 0]
                full = torch.zeros(self.B * self.S, dtype=torch.float64)
                full[keep] = mu
                y = integrate(full.view(self.B, self.S), self.dist, self.l_max)
                out.append(torch.mean((y - tgt) ** 2))
            ref = model_forward(self.spec, self.cfg, self.p, self.x, corners=self.corners)
            assert torch.equal(om.forward(om.split(self.p[:n_mlp], 32, 64, 2, 16), f, "ReLU", "Softplus", 1)[:, 0],
                               ref)
            return torch.stack(out)

    def central_differences(self, which: np.ndarray, h: float = 1e-3) -> np.ndarray:
        import torch

        out = np.empty(which.size)
        p = self.p
        n_mlp = self.cfg["n_params"]
        mlp_part = which < n_mlp
        if mlp_part.any():
            # MLP weights: features are unchanged, so only the network is re-evaluated per perturbation
            base = p[:n_mlp]
            cols = torch.from_numpy(which[mlp_part])
            plus = base.repeat(cols.numel(), 1)
            minus = plus.clone()
            ar = torch.arange(cols.numel())
            plus[ar, cols] += h
            minus[ar, cols] -= h
            out[mlp_part] = ((self._mlp_losses(plus) - self._mlp_losses(minus)) / (2 * h)).numpy()
        for k in np.flatnonzero(~mlp_part):
            i = which[k]
            v = float(p[i])
            p[i] = v + h
            lp = self.loss()
            p[i] = v - h
            lm = self.loss()
            p[i] = v
            out[k] = (lp - lm) / (2 * h)
        return out

    def fd_indices(self, seed: int = 1, n_entries: int = 256) -> np.ndarray:
        rng = np.random.default_rng(seed)
        touched = self.touched_entries()
        pick = rng.choice(touched, size=min(n_entries, touched.size), replace=False)
        return np.concatenate([np.arange(self.cfg["n_params"]), np.sort(pick)])


def rounding_edge_points(spec, n_per_level: int = 64, seed: int = 0) -> np.ndarray:
    """Points [m, D] (float32, in [0, 1]) whose position scale*x + 0.5 on some level lies within a few float32
    ulps of a cell boundary (where the floor, and so the corner indices, are most easily got wrong).  Each
    point has one such coordinate; the others are random."""
    rng = np.random.default_rng(seed)
    out = []
    for l in range(spec.n_levels):
        s = np.float32(spec.scale[l])
        k = rng.integers(1, int(spec.resolution[l]), n_per_level)
        x0 = ((k - 0.5) / float(s)).astype(np.float32)
        step = rng.integers(-3, 4, n_per_level)
        x = x0.copy()
        for i in range(n_per_level):
            for _ in range(abs(int(step[i]))):
                x[i] = np.nextafter(x[i], np.float32(np.inf) if step[i] > 0 else np.float32(-np.inf))
        x = np.clip(x, np.float32(0.0), np.float32(1.0))
        pts = rng.random((n_per_level, spec.n_dims), dtype=np.float32)
        axis = rng.integers(0, spec.n_dims, n_per_level)
        pts[np.arange(n_per_level), axis] = x
        out.append(pts)
    return np.concatenate(out).astype(np.float32)
