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
