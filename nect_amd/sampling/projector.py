"""``Projector``: the ray-batch front end of the trainer (nect/sampling/projector.py:35-131) on nect_sample.

Same constructor, ``update``/``update_angle``/``set_geometry``/``to_dict`` and ``__call__(batch_num, proj)``:
one random pixel permutation per ``update`` (torch.randperm on the device), ``points_per_batch // S`` rays per
batch with the reference's tail rule (a first batch larger than the detector takes all pixels, later
incomplete batches return (None, None)), points clamped to [0, 1] and the targets gathered through the same
permutation.  ``distances`` holds the last batch's per-sample lengths.
"""
from __future__ import annotations

import torch

from . import ct_sampling
from .geometry import Geometry


def _cuda_device(device: torch.device | str | int) -> torch.device:
    """projector.py:9-32: an int, "cuda", "cuda:k" or a torch.device naming an existing GPU."""
    if not torch.cuda.is_available():
        raise ValueError("CUDA is not available")
    max_id = torch.cuda.device_count() - 1
    if isinstance(device, int):
        if device > max_id:
            raise ValueError(f"GPU ID {device} is greater than the maximum available GPU ID {max_id}")
        return torch.device("cuda", device)
    if isinstance(device, str):
        if device == "cuda":
            return torch.device("cuda", 0)
        if not device.startswith("cuda:"):
            raise ValueError(f"'{device}' is not supported")
        device = torch.device(device)
    if device.type != "cuda":
        raise ValueError(f"'{device}' is not supported")
    index = 0 if device.index is None else device.index
    if index > max_id:
        raise ValueError(f"GPU ID {index} is greater than the maximum available GPU ID {max_id}")
    return torch.device("cuda", index)


class Projector:
    def __init__(self, geometry: Geometry, points_per_batch: int, points_per_ray: int,
                 device: torch.device | str | int, random_offset_detector: float = 0.0,
                 uniform_ray_spacing: bool = True):
        self.device = _cuda_device(device)
        if points_per_ray < 1:
            raise ValueError(f"points_per_ray ({points_per_ray}) must be greater than or equal to 1")
        if points_per_batch < points_per_ray:
            raise ValueError(f"points_per_batch ({points_per_batch}) must be greater than or equal to "
                             f"points_per_ray ({points_per_ray})")
        self.points_per_batch = points_per_batch
        self.points_per_ray = points_per_ray
        self.geometry = geometry
        self.random_offset_detector = random_offset_detector
        self.uniform_ray_spacing = uniform_ray_spacing
        self.c_geometry = self.geometry.get_c_geometry()
        self.distances = None

    def set_geometry(self, geometry: Geometry):
        self.geometry = geometry
        self.c_geometry = self.geometry.get_c_geometry()

    def to_dict(self) -> dict:
        return {"geometry": self.geometry.to_dict(), "points_per_batch": self.points_per_batch,
                "reconstruction_mode": self.geometry.reconstruction_mode, "points_per_ray": self.points_per_ray,
                "uniform_ray_spacing": self.uniform_ray_spacing}

    def update(self, angle: float, detector_binning: int | None = None, points_per_ray: int | None = None,
               random_offset_detector: float | None = None, uniform_ray_spacing: bool | None = None):
        """New view: binning, sample count and sampling options, and a fresh pixel permutation."""
        if detector_binning is not None:
            self.geometry.set_detector_binning(detector_binning)
            self.c_geometry = self.geometry.get_c_geometry()
        if points_per_ray is not None:
            self.points_per_ray = points_per_ray
        if random_offset_detector is not None:
            self.random_offset_detector = random_offset_detector
        if uniform_ray_spacing is not None:
            self.uniform_ray_spacing = uniform_ray_spacing
        b = self.geometry.detector_binning
        self.total_detector_pixels = (self.geometry.nDetector[0] // b) * (self.geometry.nDetector[1] // b)
        self.batch_size = self.points_per_batch // self.points_per_ray
        self.random_indexes = torch.randperm(self.total_detector_pixels, dtype=torch.int64, device=self.device)
        self.angle = angle
        self.distance_between_points = self.geometry.max_distance_traveled / self.points_per_ray
        self.batch_per_epoch = self.total_detector_pixels // self.batch_size
        self.batch_per_epoch += 1 if self.total_detector_pixels % self.batch_size != 0 else 0

    def update_angle(self, angle: float):
        self.angle = angle

    def __call__(self, batch_num: int, proj: torch.Tensor):
        """(points [B, S, 3] in [0, 1], targets [B]) of batch ``batch_num`` of the current view."""
        start = batch_num * self.batch_size
        batch_size = self.batch_size
        if (batch_num + 1) * batch_size > self.random_indexes.size(0):
            if batch_num == 0:
                batch_size = self.random_indexes.size(0) % batch_size
            else:
                return None, None
        points, distances = ct_sampling.sample(
            random_ray_index=self.random_indexes, geometry=self.c_geometry, angle_rad=self.angle,
            num_points_per_ray=self.points_per_ray, num_rays=batch_size, starting_ray_index=start,
            max_ray_distance_per_point=self.distance_between_points, uniform_ray_spacing=self.uniform_ray_spacing,
            random_detector_offset=self.random_offset_detector, device=self.device)
        self.distances = distances
        return points, proj[self.random_indexes[start:start + batch_size]]
