"""ORACLE (test infrastructure): float64 analytic line integrals through ellipsoid phantoms.

A phantom is a sum of ellipsoids k with attenuation mu_k (per mm), centre c_k,
semi-axes a_k and rotation R_k (columns = the ellipsoid's axes in the
volume frame).  The line integral along the segment p(s) = P1 + s (P2 - P1),
s in [0, 1], is  sum_k mu_k * |P2 - P1| * (length of {s : p(s) inside k}):
in the ellipsoid's unit-sphere frame the chord solves a quadratic.

Ray geometry follows the sampler (nect/sampling/ct_sampling.cu:226-389):
source at (-DSO, 0, 0) (cone) and detector pixel (row, col) at
x = DSD - DSO, y = -u, z = v, u/v the detector coordinates of the pixel
centre, both points rotated about z by the view angle; parallel beam puts the
source at (-x_d, y, z).  Coordinates are millimetres with the volume centred
on the origin; the model's normalised coordinate is mm / sVoxel + 0.5.

Parity pins (tests/test_oracle_analytic.py): a chord through a sphere's
centre equals its diameter; random rays agree with dense numerical
quadrature of the indicator; rotation invariance of a sphere.
"""
from __future__ import annotations

import numpy as np


def segment_integral(P1: np.ndarray, P2: np.ndarray, centres, axes, rots, mus) -> np.ndarray:
    """Integral of the phantom along segments P1[n,3] -> P2[n,3] (x, y, z mm), float64 [n]."""
    P1 = np.asarray(P1, np.float64)
    P2 = np.asarray(P2, np.float64)
    d = P2 - P1
    seg = np.linalg.norm(d, axis=1)
    out = np.zeros(P1.shape[0])
    for c, a, R, mu in zip(centres, axes, rots, mus):
        q = ((P1 - c) @ R) / a          # in the unit-sphere frame
        e = (d @ R) / a
        A = np.sum(e * e, 1)
        B = 2 * np.sum(q * e, 1)
        C = np.sum(q * q, 1) - 1
        disc = B * B - 4 * A * C
        ok = disc > 0
        sq = np.sqrt(np.where(ok, disc, 0))
        s0 = np.clip((-B - sq) / (2 * A), 0, 1)
        s1 = np.clip((-B + sq) / (2 * A), 0, 1)
        out += np.where(ok, mu * seg * np.maximum(s1 - s0, 0), 0)
    return out


def ray_points(mode: str, n_det, d_det, dso, dsd, angle: float, pixels: np.ndarray, reach: float = 0.0):
    """(P1, P2) of detector pixels (index row*cols + col) at pixel centres, x/y/z mm."""
    rows, cols = n_det
    col = pixels % cols
    row = pixels // cols
    u = (col - (cols - 1) / 2) * d_det[1]
    v = (row - (rows - 1) / 2) * d_det[0]
    n = pixels.size
    if mode == "cone":
        p1 = np.tile([-dso, 0.0, 0.0], (n, 1))
        p2 = np.stack([np.full(n, dsd - dso), -u, v], 1)
    else:
        p2 = np.stack([np.full(n, reach), -u, v], 1)
        p1 = np.stack([np.full(n, -reach), -u, v], 1)
    c, s = np.cos(angle), np.sin(angle)
    Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    return p1 @ Rz.T, p2 @ Rz.T


def voxel_value(points: np.ndarray, centres, axes, rots, mus) -> np.ndarray:
    """Attenuation at points [n, 3] (x, y, z mm)."""
    out = np.zeros(points.shape[0])
    for c, a, R, mu in zip(centres, axes, rots, mus):
        q = ((points - c) @ R) / a
        out += np.where(np.sum(q * q, 1) <= 1, mu, 0)
    return out
