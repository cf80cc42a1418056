"""``Geometry``: the acquisition geometry object of nect/sampling/geometry.py:26-439.

Same constructor keywords, derived sizes (``sVoxel``, ``sDetector``), the
sample-spacing normaliser ``max_distance_traveled`` (geometry.py:156-174),
YAML/dict round trips, and ``get_c_geometry()`` mapping onto the sampler's
geometry structs with detector binning (geometry.py:322-407).
"""
from __future__ import annotations

from pathlib import Path
from warnings import warn

import numpy as np
import torch
import yaml

from . import ct_sampling


class Geometry:
    def __init__(self, nDetector, dDetector, mode: str, DSD: float | None = None, DSO: float | None = None,
                 nVoxel=None, dVoxel=None, radius: float | None = None, height: float | None = None,
                 offOrigin=(0.0, 0.0, 0.0), COR: float = 0.0, offDetector=(0.0, 0.0), rotDetector=(0.0, 0.0, 0.0),
                 reconstruction_mode: str = "voxel", detector_binning: int = 1, angles=None, radians: bool = True,
                 timesteps=None):
        if mode not in ["cone", "parallel"]:
            raise ValueError(f"geometry mode must be 'cone' or 'parallel', got '{mode}'")
        if mode == "cone":
            if DSD is None:
                raise ValueError("a cone geometry needs DSD (source to detector distance)")
            if DSO is None:
                raise ValueError("a cone geometry needs DSO (source to origin distance)")
            self.DSD = float(DSD)
            self.DSO = float(DSO)
        else:
            if DSD is not None:
                warn("a parallel geometry has no source: DSD is ignored")
            if DSO is not None:
                warn("a parallel geometry has no source: DSO is ignored")
            self.DSD = None
            self.DSO = None
        if reconstruction_mode not in ["voxel", "cylindrical"]:
            raise ValueError(f"reconstruction_mode must be 'voxel' or 'cylindrical', got '{reconstruction_mode}'")
        self.mode = mode
        self.reconstruction_mode = reconstruction_mode
        self.detector_binning = detector_binning
        self.dDetector = self._floats(dDetector, 2, "dDetector")
        self.nDetector = self._ints(nDetector, 2, "nDetector")
        self.offDetector = self._floats(offDetector, 2, "offDetector")
        self.rotDetector = self._floats(rotDetector, 3, "rotDetector")
        self.COR = COR
        self.set_angles(angles, radians)
        self.set_timesteps(timesteps)
        voxels_given = nVoxel is not None and dVoxel is not None
        if radius is None or height is None:
            self.nVoxel = self._ints(nVoxel, 3, "nVoxel")
            self.dVoxel = self._floats(dVoxel, 3, "dVoxel")
            self.sVoxel = tuple(self.nVoxel[i] * self.dVoxel[i] for i in range(3))
            self.offOrigin = self._floats(offOrigin, 3, "offOrigin")
            if self.reconstruction_mode == "cylindrical":
                self.radius = radius if radius is not None else max(self.sVoxel[1], self.sVoxel[2]) / 2
                self.height = height if height is not None else self.sVoxel[0]
            else:
                self.radius = None
                self.height = None
        else:
            self.radius = radius
            self.height = height
            self.offOrigin = self._floats(offOrigin, 3, "offOrigin")
            self.nVoxel = self._ints(nVoxel, 3, "nVoxel") if nVoxel is not None else (
                self.nDetector[0], self.nDetector[1], self.nDetector[1])
            self.dVoxel = self._floats(dVoxel, 3, "dVoxel") if dVoxel is not None else (
                self.nDetector[0] / height, self.nDetector[1] / (2 * radius), self.nDetector[1] / (2 * radius))
            self.sVoxel = tuple(self.nVoxel[i] * self.dVoxel[i] for i in range(3))
        s, d = self.sVoxel, self.dVoxel
        if self.mode == "cone":
            if not voxels_given and reconstruction_mode == "cylindrical":
                max_length = self.radius * 2
                theta = np.arctan((self.height / 2) / (self.DSO + max_length / 2))
            else:
                max_length = ((s[1] - d[1]) ** 2 + (s[2] - d[2]) ** 2) ** 0.5
                theta = np.arctan(((s[0] - d[0]) / 2) / (self.DSO + max_length / 2))
            self.max_distance_traveled = float(max_length / np.cos(theta))
        else:
            if self.reconstruction_mode == "voxel":
                self.max_distance_traveled = float(((s[1] - d[1]) ** 2 + (s[2] - d[2]) ** 2) ** 0.5)
            elif not voxels_given:
                self.max_distance_traveled = float(self.radius * 2)
            else:
                self.max_distance_traveled = float(max(s[1], s[2]))
        self.sDetector = (self.nDetector[0] * self.dDetector[0], self.nDetector[1] * self.dDetector[1])

    # ---------------------------------------------------------------- I/O

    @classmethod
    def from_yaml(cls, path: str | Path, reconstruction_mode: str | None = None) -> "Geometry":
        with open(path, "r") as f:
            cfg = yaml.safe_load(f)
        return cls.from_dict(cfg, reconstruction_mode)

    @classmethod
    def from_dict(cls, cfg: dict, reconstruction_mode: str | None = None) -> "Geometry":
        if reconstruction_mode is None:
            reconstruction_mode = cfg.get("reconstruction_mode", "voxel")
        return cls(nDetector=cfg["nDetector"], dDetector=cfg["dDetector"], mode=cfg["mode"], DSD=cfg.get("DSD"),
                   DSO=cfg.get("DSO"), nVoxel=cfg.get("nVoxel"), dVoxel=cfg.get("dVoxel"), radius=cfg.get("radius"),
                   height=cfg.get("height"), offOrigin=cfg.get("offOrigin", (0.0, 0.0, 0.0)), COR=cfg.get("COR", 0.0),
                   offDetector=cfg.get("offDetector", (0.0, 0.0)), rotDetector=cfg.get("rotDetector", (0.0, 0.0, 0.0)),
                   reconstruction_mode=reconstruction_mode, detector_binning=1, angles=cfg.get("angles"),
                   timesteps=cfg.get("timesteps"), radians=cfg.get("radians", True))

    def to_dict(self) -> dict:
        return {
            "nDetector": list(self.nDetector), "dDetector": list(self.dDetector), "sDetector": list(self.sDetector),
            "mode": self.mode, "DSD": self.DSD, "DSO": self.DSO, "nVoxel": list(self.nVoxel),
            "dVoxel": list(self.dVoxel), "sVoxel": list(self.sVoxel), "radius": self.radius, "height": self.height,
            "angles": None if self.angles is None else [float(a) for a in np.asarray(self.angles)],
            "radians": True, "offOrigin": list(self.offOrigin), "COR": self.COR,
            "offDetector": list(self.offDetector), "rotDetector": list(self.rotDetector),
            "timesteps": None if self.timesteps is None else [float(t) for t in np.asarray(self.timesteps)],
        }

    def set_angles(self, angles, radians: bool = True):
        self.angles = angles
        if self.angles is not None:
            if isinstance(self.angles, torch.Tensor):
                self.angles = self.angles.cpu().numpy()
            self.angles = np.asarray(self.angles, dtype=np.float64)
            if radians is False:
                self.angles = np.radians(self.angles)

    def set_detector_binning(self, detector_binning: int):
        self.detector_binning = detector_binning

    def set_timesteps(self, timesteps):
        self.timesteps = timesteps

    @property
    def num_angles(self) -> int:
        return 0 if self.angles is None else len(self.angles)

    # -------------------------------------------------------- sampler structs

    def get_c_geometry(self):
        b = self.detector_binning
        n_det = (self.nDetector[0] // b, self.nDetector[1] // b)
        d_det = (self.dDetector[0] * b, self.dDetector[1] * b)
        common = dict(nDetector=n_det, dDetector=d_det, offOrigin=self.offOrigin, offDetector=self.offDetector,
                      rotDetector=self.rotDetector, COR=self.COR)
        if self.mode == "cone":
            if self.DSD <= self.DSO:
                raise ValueError(f"the detector must lie beyond the origin: DSD ({self.DSD}) <= DSO ({self.DSO})")
            if self.DSO <= 0:
                raise ValueError(f"the source must lie before the origin: DSO ({self.DSO}) <= 0")
            if self.reconstruction_mode == "cylindrical":
                return ct_sampling.ConeGeometryCylindrical(self.DSO, self.DSD - self.DSO, object_radius=self.radius,
                                                           object_height=self.height, **common)
            return ct_sampling.ConeGeometryVoxel(self.DSO, self.DSD - self.DSO, sVoxel=self.sVoxel,
                                                 dVoxel=self.dVoxel, **common)
        if self.reconstruction_mode == "cylindrical":
            return ct_sampling.ParallelGeometryCylindrical(object_radius=self.radius, object_height=self.height,
                                                           **common)
        return ct_sampling.ParallelGeometryVoxel(sVoxel=self.sVoxel, dVoxel=self.dVoxel, **common)

    # --------------------------------------------------------------- checks

    @staticmethod
    def _tuple(t, length: int, name: str):
        if t is None:
            raise ValueError(f"the geometry needs {name}")
        if isinstance(t, (np.ndarray, torch.Tensor)):
            if tuple(t.shape) != (length,):
                raise ValueError(f"{name} takes {length} values")
        elif len(t) != length:
            raise ValueError(f"{name} takes {length} values")
        return tuple(t[i] for i in range(length))

    @classmethod
    def _floats(cls, t, length: int, name: str):
        t = cls._tuple(t, length, name)
        out = []
        for i, v in enumerate(t):
            if isinstance(v, (torch.Tensor, np.generic)):
                v = v.item()
            if not isinstance(v, (float, int)):
                raise ValueError(f"{name}[{i}] must be a number")
            out.append(float(v))
        return tuple(out)

    @classmethod
    def _ints(cls, t, length: int, name: str):
        # every element is checked (the reference returns after checking only the first,
        # nect/sampling/geometry.py:416-425)
        t = cls._tuple(t, length, name)
        out = []
        for i, v in enumerate(t):
            if isinstance(v, (torch.Tensor, np.generic)):
                v = v.item()
            if isinstance(v, float):
                if not v.is_integer():
                    raise ValueError(f"{name}[{i}] should be an integer")
            elif not isinstance(v, int):
                raise ValueError(f"{name}[{i}] should be an integer")
            out.append(int(v))
        return tuple(out)
