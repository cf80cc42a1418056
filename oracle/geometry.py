"""ORACLE (test infrastructure): geometry bookkeeping of the reference.

Restates nect/sampling/geometry.py:91-178 (derived sizes and the sample-spacing
normaliser ``max_distance_traveled``) and :322-407 (mapping onto the sampler's
C structs with detector binning) as plain Python on float64.

Parity: pinned by closed forms in tests/test_oracle_sampler.py (hand-computed
L_max of the 64^3 cone configuration, 44.568 mm, and of the parallel case).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field


@dataclass
class OracleGeometry:
    mode: str                      # "cone" | "parallel"
    nDetector: tuple[int, int]
    dDetector: tuple[float, float]
    DSD: float | None = None
    DSO: float | None = None
    nVoxel: tuple[int, int, int] | None = None
    dVoxel: tuple[float, float, float] | None = None
    radius: float | None = None
    height: float | None = None
    COR: float = 0.0
    offDetector: tuple[float, float] = (0.0, 0.0)
    rotDetector: tuple[float, float, float] = (0.0, 0.0, 0.0)
    reconstruction_mode: str = "voxel"
    detector_binning: int = 1
    sVoxel: tuple[float, float, float] = field(init=False)
    max_distance_traveled: float = field(init=False)

    def __post_init__(self):
        voxels_given = self.nVoxel is not None and self.dVoxel is not None
        # geometry.py:119-155
        if self.radius is None or self.height is None:
            self.sVoxel = tuple(self.nVoxel[i] * self.dVoxel[i] for i in range(3))
            if self.reconstruction_mode == "cylindrical":
                if self.radius is None:
                    self.radius = max(self.sVoxel[1], self.sVoxel[2]) / 2
                if self.height is None:
                    self.height = self.sVoxel[0]
        else:
            if self.nVoxel is None:
                self.nVoxel = (self.nDetector[0], self.nDetector[1], self.nDetector[1])
            if self.dVoxel is None:
                self.dVoxel = (self.nDetector[0] / self.height, self.nDetector[1] / (2 * self.radius),
                               self.nDetector[1] / (2 * self.radius))
            self.sVoxel = tuple(self.nVoxel[i] * self.dVoxel[i] for i in range(3))
        # geometry.py:156-174
        s, d = self.sVoxel, self.dVoxel
        if self.mode == "cone":
            if not voxels_given and self.reconstruction_mode == "cylindrical":
                max_length = self.radius * 2
                theta = math.atan((self.height / 2) / (self.DSO + max_length / 2))
            else:
                max_length = ((s[1] - d[1]) ** 2 + (s[2] - d[2]) ** 2) ** 0.5
                theta = math.atan(((s[0] - d[0]) / 2) / (self.DSO + max_length / 2))
            self.max_distance_traveled = max_length / math.cos(theta)
        else:
            if self.reconstruction_mode == "voxel":
                self.max_distance_traveled = ((s[1] - d[1]) ** 2 + (s[2] - d[2]) ** 2) ** 0.5
            elif not voxels_given:
                self.max_distance_traveled = self.radius * 2
            else:
                self.max_distance_traveled = max(s[1], s[2])

    def c_struct_fields(self) -> dict:
        """Fields of the sampler struct (geometry.py:322-407), before float32 rounding."""
        b = self.detector_binning
        n_det = (self.nDetector[0] // b, self.nDetector[1] // b)
        d_det = (self.dDetector[0] * b, self.dDetector[1] * b)
        cyl = self.reconstruction_mode == "cylindrical"
        if self.mode == "cone":
            kind = 2 if cyl else 0
        else:
            kind = 3 if cyl else 1
        return {
            "kind": kind,
            "n_det": n_det,
            "d_det": d_det,
            "off_det": tuple(self.offDetector),
            "rot_det": tuple(self.rotDetector),
            "cor": self.COR,
            "dso": self.DSO if self.mode == "cone" else 0.0,
            "dod": (self.DSD - self.DSO) if self.mode == "cone" else 0.0,
            "s_voxel": tuple(self.sVoxel),
            "d_voxel": tuple(self.dVoxel) if self.dVoxel is not None else (0.0, 0.0, 0.0),
            "radius": self.radius if cyl else 0.0,
            "height": self.height if cyl else 0.0,
            "remove_top": 0.0,
            "remove_bottom": 0.0,
        }
