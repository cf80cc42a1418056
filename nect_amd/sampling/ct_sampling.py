"""Drop-in for the reference's compiled ``nect.sampling.ct_sampling`` module.

Mirrors nect/sampling/ct_sampling.cu:650-808: four geometry classes with the
same constructor keywords and read/write fields, and ``sample(...)`` with the
same keyword signature, returning ``(points [num_rays, S, 3], distances
[num_rays])`` on the GPU.  The work runs in the gfx950 kernel behind
``nect_sample`` (include/nect_amd.h).

Differences, all additive:
  * ``jitter`` / ``detector_offset`` may be passed in; when omitted they are
    drawn on the device exactly as the reference draws them
    (``rand - 0.5`` and ``(rand - 0.5) * random_detector_offset``,
    ct_sampling.cu:619-622).
  * ``return_counts=True`` also returns the per-ray sample count.
  * The call is asynchronous on torch's current stream (the reference
    synchronises the device after every call, ct_sampling.cu:640).
"""
from __future__ import annotations

from typing import Sequence

import torch

from .. import _lib

CONE_VOXEL = 0
PARALLEL_VOXEL = 1
CONE_CYLINDRICAL = 2
PARALLEL_CYLINDRICAL = 3


def _pair(v, name):
    v = tuple(v)
    if len(v) != 2:
        raise ValueError(f"{name} must have 2 elements")
    return v


def _triple(v, name):
    v = tuple(v)
    if len(v) != 3:
        raise ValueError(f"{name} must have 3 elements")
    return v


class _BaseGeometry:
    kind: int

    def __init__(self, nDetector, dDetector, offOrigin=(0.0, 0.0, 0.0), offDetector=(0.0, 0.0),
                 rotDetector=(0.0, 0.0, 0.0), COR=0.0):
        self.nDetector = tuple(int(x) for x in _pair(nDetector, "nDetector"))
        self.dDetector = tuple(float(x) for x in _pair(dDetector, "dDetector"))
        # offOrigin is stored but, as in the reference, never used by the sampler.
        self.offOrigin = tuple(float(x) for x in _triple(offOrigin, "offOrigin"))
        self.offDetector = tuple(float(x) for x in _pair(offDetector, "offDetector"))
        self.rotDetector = tuple(float(x) for x in _triple(rotDetector, "rotDetector"))
        self.COR = float(COR)

    def _struct(self) -> _lib.NectGeometry:
        g = _lib.NectGeometry()
        g.kind = self.kind
        g.n_det[0], g.n_det[1] = self.nDetector
        g.d_det[0], g.d_det[1] = self.dDetector
        g.off_det[0], g.off_det[1] = self.offDetector
        for i in range(3):
            g.rot_det[i] = self.rotDetector[i]
        g.cor = self.COR
        dso = getattr(self, "distance_source_to_origin", 0.0)
        dod = getattr(self, "distance_origin_to_detector", 0.0)
        g.dso, g.dod = dso, dod
        sv = getattr(self, "sVoxel", (0.0, 0.0, 0.0))
        dv = getattr(self, "dVoxel", (0.0, 0.0, 0.0))
        for i in range(3):
            g.s_voxel[i] = sv[i]
            g.d_voxel[i] = dv[i]
        g.radius = getattr(self, "object_radius", 0.0)
        g.height = getattr(self, "object_height", 0.0)
        g.remove_top = getattr(self, "remove_factor_top", 0.0)
        g.remove_bottom = getattr(self, "remove_factor_bottom", 0.0)
        return g


class ConeGeometryVoxel(_BaseGeometry):
    """ct_sampling.cu:73-86 / bindings :705-728."""

    kind = CONE_VOXEL

    def __init__(self, distance_source_to_origin, distance_origin_to_detector, nDetector, dDetector,
                 offOrigin=(0.0, 0.0, 0.0), offDetector=(0.0, 0.0), rotDetector=(0.0, 0.0, 0.0), COR=0.0, *,
                 sVoxel, dVoxel):
        super().__init__(nDetector, dDetector, offOrigin, offDetector, rotDetector, COR)
        self.distance_source_to_origin = float(distance_source_to_origin)
        self.distance_origin_to_detector = float(distance_origin_to_detector)
        self.sVoxel = tuple(float(x) for x in _triple(sVoxel, "sVoxel"))
        self.dVoxel = tuple(float(x) for x in _triple(dVoxel, "dVoxel"))


class ParallelGeometryVoxel(_BaseGeometry):
    """ct_sampling.cu:88-98 / bindings :730-749."""

    kind = PARALLEL_VOXEL

    def __init__(self, nDetector, dDetector, offOrigin=(0.0, 0.0, 0.0), offDetector=(0.0, 0.0),
                 rotDetector=(0.0, 0.0, 0.0), COR=0.0, *, sVoxel, dVoxel):
        super().__init__(nDetector, dDetector, offOrigin, offDetector, rotDetector, COR)
        self.sVoxel = tuple(float(x) for x in _triple(sVoxel, "sVoxel"))
        self.dVoxel = tuple(float(x) for x in _triple(dVoxel, "dVoxel"))


class ConeGeometryCylindrical(_BaseGeometry):
    """ct_sampling.cu:100-115 / bindings :676-703."""

    kind = CONE_CYLINDRICAL

    def __init__(self, distance_source_to_origin, distance_origin_to_detector, nDetector, dDetector,
                 offOrigin=(0.0, 0.0, 0.0), offDetector=(0.0, 0.0), rotDetector=(0.0, 0.0, 0.0), COR=0.0, *,
                 object_radius, object_height, remove_factor_top=0.0, remove_factor_bottom=0.0):
        super().__init__(nDetector, dDetector, offOrigin, offDetector, rotDetector, COR)
        self.distance_source_to_origin = float(distance_source_to_origin)
        self.distance_origin_to_detector = float(distance_origin_to_detector)
        self.object_radius = float(object_radius)
        self.object_height = float(object_height)
        self.remove_factor_top = float(remove_factor_top)
        self.remove_factor_bottom = float(remove_factor_bottom)


class ParallelGeometryCylindrical(_BaseGeometry):
    """ct_sampling.cu:54-66 / bindings :650-673."""

    kind = PARALLEL_CYLINDRICAL

    def __init__(self, nDetector, dDetector, object_radius, object_height, offOrigin=(0.0, 0.0, 0.0),
                 offDetector=(0.0, 0.0), rotDetector=(0.0, 0.0, 0.0), COR=0.0, remove_factor_top=0.0,
                 remove_factor_bottom=0.0):
        super().__init__(nDetector, dDetector, offOrigin, offDetector, rotDetector, COR)
        self.object_radius = float(object_radius)
        self.object_height = float(object_height)
        self.remove_factor_top = float(remove_factor_top)
        self.remove_factor_bottom = float(remove_factor_bottom)


GeometryStruct = ConeGeometryVoxel | ParallelGeometryVoxel | ConeGeometryCylindrical | ParallelGeometryCylindrical


def _device_of(device) -> torch.device:
    if isinstance(device, int):
        return torch.device("cuda", device)
    d = torch.device(device)
    if d.type != "cuda":
        raise ValueError(f"sample() runs on the GPU only, got device {d}")
    return torch.device("cuda", d.index if d.index is not None else torch.cuda.current_device())


def draw_randomness(num_rays: int, random_detector_offset: float, device, generator=None):
    """(jitter [B], detector_offset [B, 2]) drawn as ct_sampling.cu:619-622 draws them."""
    jitter = torch.rand(num_rays, device=device, generator=generator) - 0.5
    offset = (torch.rand(num_rays, 2, device=device, generator=generator) - 0.5) * random_detector_offset
    return jitter, offset


def sample(random_ray_index: torch.Tensor, geometry: GeometryStruct, angle_rad: float, num_points_per_ray: int,
           num_rays: int, starting_ray_index: int, max_ray_distance_per_point: float, uniform_ray_spacing: bool,
           random_detector_offset: float, device, *, jitter: torch.Tensor | None = None,
           detector_offset: torch.Tensor | None = None, return_counts: bool = False):
    """Samples points along ``num_rays`` detector rays of one view (ct_sampling.cu:606-648).

    Returns ``(points, distances)`` (plus ``counts`` if requested); ``points`` is
    [num_rays, S, 3] float32 (z, y, x), zero in unwritten slots.  Unlike the
    reference the points are already clamped to [0, 1] (projector.py:127 does
    it in place right after the call, so callers see the same values).
    """
    dev = _device_of(device)
    if random_ray_index.dtype != torch.int64 or random_ray_index.dim() != 1:
        raise ValueError("random_ray_index must be a 1-D int64 tensor")
    if random_ray_index.device != dev:
        raise ValueError("random_ray_index must live on the sampling device")
    num_rays = int(num_rays)
    S = int(num_points_per_ray)
    if num_rays < 0:
        raise ValueError("num_rays must be >= 0")
    if S < 1:
        raise ValueError("num_points_per_ray must be >= 1")
    if jitter is None or detector_offset is None:
        j, o = draw_randomness(num_rays, random_detector_offset, dev)
        jitter = j if jitter is None else jitter
        detector_offset = o if detector_offset is None else detector_offset
    jitter = jitter.to(device=dev, dtype=torch.float32).contiguous()
    detector_offset = detector_offset.to(device=dev, dtype=torch.float32).contiguous()
    if jitter.shape != (num_rays,) or detector_offset.shape != (num_rays, 2):
        raise ValueError("jitter must be [num_rays] and detector_offset [num_rays, 2]")
    points = torch.empty((num_rays, S, 3), dtype=torch.float32, device=dev)
    distances = torch.empty((num_rays,), dtype=torch.float32, device=dev)
    counts = torch.empty((num_rays,), dtype=torch.int32, device=dev)
    ray_index = random_ray_index.contiguous()
    g = geometry._struct()
    if num_rays > 0:
        _lib.call("nect_sample", g, float(angle_rad), _lib.ptr(ray_index), ray_index.numel(), num_rays,
                  int(starting_ray_index), S, float(max_ray_distance_per_point), int(bool(uniform_ray_spacing)),
                  _lib.ptr(jitter), _lib.ptr(detector_offset), _lib.ptr(points), _lib.ptr(distances),
                  _lib.ptr(counts), _lib.stream_of(dev))
    if return_counts:
        return points, distances, counts
    return points, distances


def ray_setup(random_ray_index: torch.Tensor, geometry: GeometryStruct, angle_rad: float, num_points_per_ray: int,
              num_rays: int, starting_ray_index: int, max_ray_distance_per_point: float, uniform_ray_spacing: bool,
              jitter: torch.Tensor, detector_offset: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """Fused-path ray descriptors: uint8 tensor [num_rays, 64] viewing NectRay records."""
    dev = random_ray_index.device
    if out is None:
        out = torch.empty((num_rays, _lib.RAY_BYTES), dtype=torch.uint8, device=dev)
    g = geometry._struct()
    _lib.call("nect_ray_setup", g, float(angle_rad), _lib.ptr(random_ray_index), random_ray_index.numel(),
              int(num_rays), int(starting_ray_index), int(num_points_per_ray), float(max_ray_distance_per_point),
              int(bool(uniform_ray_spacing)), _lib.ptr(jitter), _lib.ptr(detector_offset), _lib.ptr(out),
              _lib.stream_of(dev))
    return out


def ray_fields(rays: torch.Tensor):
    """Views (start [B,3] f64 (x,y,z), step [B,3] f64, count [B] i32, distance [B] f32) of descriptors."""
    b = rays.shape[0]
    d = rays.view(torch.float64).view(b, 8)
    i = rays.view(torch.int32).view(b, 16)
    f = rays.view(torch.float32).view(b, 16)
    return d[:, 0:3], d[:, 3:6], i[:, 12], f[:, 13]


def ray_tiles(rays: torch.Tensor, max_tiles: int):
    """(tile_offsets [B+1], tile_ray [max_tiles], n_tiles [1]) for descriptors ``rays``."""
    dev = rays.device
    b = rays.shape[0]
    tile_offsets = torch.empty(b + 1, dtype=torch.int32, device=dev)
    tile_ray = torch.empty(max_tiles, dtype=torch.int32, device=dev)
    n_tiles = torch.empty(1, dtype=torch.int32, device=dev)
    _lib.call("nect_ray_tiles", _lib.ptr(rays), b, int(max_tiles), _lib.ptr(tile_offsets), _lib.ptr(tile_ray),
              _lib.ptr(n_tiles), _lib.stream_of(dev))
    return tile_offsets, tile_ray, n_tiles


def ray_positions(rays, tile_offsets, tile_ray, n_tiles, max_tiles: int):
    """Slot-ordered sample positions [max_tiles*32, 3] and validity [max_tiles*32] (int32)."""
    dev = rays.device
    pos = torch.empty((max_tiles * 32, 3), dtype=torch.float32, device=dev)
    valid = torch.empty(max_tiles * 32, dtype=torch.int32, device=dev)
    _lib.call("nect_ray_positions", _lib.ptr(rays), _lib.ptr(tile_offsets), _lib.ptr(tile_ray), _lib.ptr(n_tiles),
              int(max_tiles), _lib.ptr(pos), _lib.ptr(valid), _lib.stream_of(dev))
    return pos, valid


def tiles_needed(num_rays: int, num_points_per_ray: int) -> int:
    return num_rays * ((num_points_per_ray + 31) // 32)


__all__: Sequence[str] = [
    "ConeGeometryVoxel", "ParallelGeometryVoxel", "ConeGeometryCylindrical", "ParallelGeometryCylindrical",
    "sample", "ray_setup", "ray_fields", "ray_tiles", "ray_positions", "draw_randomness", "tiles_needed",
]
