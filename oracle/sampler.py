"""ORACLE (test infrastructure): Python wrapper of the C sampler restatement.

``oracle_sample`` restates ct_sampling.sample (nect/sampling/ct_sampling.cu:
606-648) + the projector's clamp (nect/sampling/projector.py:127) for a batch.

Parity pins (tests/test_oracle_sampler.py):
  * exact rational (fractions.Fraction) ray/box intersection on small detectors
    for several views -> entry/exit points and sample counts;
  * the closed-form count of the centre ray of an odd detector at angle 0;
  * parallel-beam rays at angle 0 run along +x with constant (z, y);
  * invariants: points in [0, 1], spacing = step per axis, count <= S, misses
    give zero points and distance 0.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"


class OrcGeometry(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("n_det", ctypes.c_int32 * 2),
        ("d_det", ctypes.c_float * 2),
        ("off_det", ctypes.c_float * 2),
        ("rot_det", ctypes.c_float * 3),
        ("cor", ctypes.c_float),
        ("dso", ctypes.c_float),
        ("dod", ctypes.c_float),
        ("s_voxel", ctypes.c_float * 3),
        ("d_voxel", ctypes.c_float * 3),
        ("radius", ctypes.c_float),
        ("height", ctypes.c_float),
        ("remove_top", ctypes.c_float),
        ("remove_bottom", ctypes.c_float),
    ]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            from . import build as _build

            _build.build()
        _lib = ctypes.CDLL(str(LIB))
        _lib.oracle_sample.restype = None
        _lib.oracle_sample.argtypes = [
            ctypes.POINTER(OrcGeometry), ctypes.c_float, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
            ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ]
        _lib.oracle_ray_entry_exit.restype = ctypes.c_int32
        _lib.oracle_ray_entry_exit.argtypes = [
            ctypes.POINTER(OrcGeometry), ctypes.c_float, ctypes.c_int32, ctypes.c_float, ctypes.c_float,
            ctypes.c_void_p,
        ]
    return _lib


def make_struct(fields: dict) -> OrcGeometry:
    g = OrcGeometry()
    g.kind = fields["kind"]
    g.n_det[0], g.n_det[1] = fields["n_det"]
    g.d_det[0], g.d_det[1] = fields["d_det"]
    g.off_det[0], g.off_det[1] = fields.get("off_det", (0.0, 0.0))
    rot = fields.get("rot_det", (0.0, 0.0, 0.0))
    for i in range(3):
        g.rot_det[i] = rot[i]
    g.cor = fields.get("cor", 0.0)
    g.dso = fields.get("dso", 0.0)
    g.dod = fields.get("dod", 0.0)
    sv = fields.get("s_voxel", (0.0, 0.0, 0.0))
    dv = fields.get("d_voxel", (0.0, 0.0, 0.0))
    for i in range(3):
        g.s_voxel[i] = sv[i]
        g.d_voxel[i] = dv[i]
    g.radius = fields.get("radius", 0.0)
    g.height = fields.get("height", 0.0)
    g.remove_top = fields.get("remove_top", 0.0)
    g.remove_bottom = fields.get("remove_bottom", 0.0)
    return g


def oracle_sample(fields: dict, angle_rad: float, ray_index: np.ndarray, num_rays: int, starting_ray_index: int,
                  S: int, step_mm: float, uniform: bool, jitter: np.ndarray, det_offset: np.ndarray):
    """Returns (points [B,S,3] f32, distances [B] f32, counts [B] i32)."""
    lib = _load()
    g = make_struct(fields)
    ray_index = np.ascontiguousarray(ray_index, dtype=np.int64)
    jitter = np.ascontiguousarray(jitter, dtype=np.float32)
    det_offset = np.ascontiguousarray(det_offset, dtype=np.float32).reshape(num_rays, 2)
    points = np.empty((num_rays, S, 3), dtype=np.float32)
    distances = np.empty(num_rays, dtype=np.float32)
    counts = np.empty(num_rays, dtype=np.int32)
    lib.oracle_sample(ctypes.byref(g), float(angle_rad), ray_index.ctypes.data, ray_index.size, num_rays,
                      starting_ray_index, S, float(step_mm), int(bool(uniform)), jitter.ctypes.data,
                      det_offset.ctypes.data, points.ctypes.data, distances.ctypes.data, counts.ctypes.data)
    return points, distances, counts


def oracle_entry_exit(fields: dict, angle_rad: float, ray: int, off_u: float = 0.0, off_v: float = 0.0):
    """(entry xyz, exit xyz) in mm as float64 arrays, or None for a miss."""
    lib = _load()
    g = make_struct(fields)
    out = np.zeros(6, dtype=np.float64)
    ok = lib.oracle_ray_entry_exit(ctypes.byref(g), float(angle_rad), int(ray), float(off_u), float(off_v),
                                   out.ctypes.data)
    if not ok:
        return None
    return out[:3].copy(), out[3:].copy()
