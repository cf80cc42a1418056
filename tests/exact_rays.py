"""High-precision restatement of ray/box sampling for pinning the oracle.

Everything the reference fixes in float32 (view cosine/sine, detector
coordinates, geometry fields, jitter, spacing, tolerance) enters as its exact
float32 value; everything after that is exact rational arithmetic, except the
ray length, taken to 50 significant digits.  The result is the mathematically
exact version of the sampler, against which the oracle's float/double pipeline
must agree everywhere except within a small window around ties.
"""
from __future__ import annotations

import math
from decimal import Decimal, getcontext
from fractions import Fraction as Fr

import numpy as np

getcontext().prec = 50

f32 = np.float32


def F(x) -> Fr:
    return Fr(float(x))


def fsqrt(q: Fr) -> Fr:
    d = (Decimal(q.numerator) / Decimal(q.denominator)).sqrt()
    return Fr(d)


def exact_voxel_ray(fields: dict, angle: float, ray: int, ou: float, ov: float, jitter: float, S: int,
                    step_mm: float):
    """Exact (start_n, step_n, end_n) as Fraction triples (x, y, z), or None on a miss."""
    cone = fields["kind"] == 0
    rows, cols = fields["n_det"]
    ang = f32(angle)
    c = F(f32(math.cos(float(ang))))
    s = F(f32(math.sin(float(ang))))
    col, row = ray % cols, ray // cols
    # detector coordinates in float32, as the reference computes them
    u = ((f32(col) - f32(f32(cols - 1) / f32(2))) + f32(ou)) * f32(fields["d_det"][1]) + f32(fields["off_det"][1])
    v = ((f32(row) - f32(f32(rows - 1) / f32(2))) + f32(ov)) * f32(fields["d_det"][0]) + f32(fields["off_det"][0])
    p2 = [Fr(0), -F(u), F(v)]
    sv = [F(f32(x)) for x in fields["s_voxel"]]  # z, y, x
    dv = [F(f32(x)) for x in fields["d_voxel"]]
    if cone:
        p2[0] += F(f32(fields["dod"]))
        p1 = [-F(f32(fields["dso"])), Fr(0), Fr(0)]
    else:
        po = f32(np.sqrt(f32(f32(sv[2]) * f32(sv[2]) + f32(sv[1]) * f32(sv[1]))))
        p2[0] += F(po)
        p1 = [-p2[0], p2[1], p2[2]]

    def rot(p):
        return [c * p[0] - s * p[1], s * p[0] + c * p[1], p[2]]

    p1, p2 = rot(p1), rot(p2)
    hx = F(f32(f32(f32(fields["s_voxel"][2]) - f32(fields["d_voxel"][2]) / f32(2)) / f32(2)))
    hy = F(f32(f32(f32(fields["s_voxel"][1]) - f32(fields["d_voxel"][1]) / f32(2)) / f32(2)))
    hz = F(f32(f32(f32(fields["s_voxel"][0]) - f32(fields["d_voxel"][0]) / f32(2)) / f32(2)))
    half = [hx, hy, hz]
    t0, t1 = Fr(-10 ** 30), Fr(10 ** 30)
    for a in range(3):
        d = p2[a] - p1[a]
        if d == 0:
            if not (-half[a] <= p1[a] <= half[a]):
                return None
            continue
        ta, tb = (-half[a] - p1[a]) / d, (half[a] - p1[a]) / d
        if ta > tb:
            ta, tb = tb, ta
        t0, t1 = max(t0, ta), min(t1, tb)
    if t0 > t1 or not (0 <= t0 <= 1) or not (0 <= t1 <= 1):
        return None
    st = [p1[a] + t0 * (p2[a] - p1[a]) for a in range(3)]
    en = [p1[a] + t1 * (p2[a] - p1[a]) for a in range(3)]
    dd = [en[a] - st[a] for a in range(3)]
    length = fsqrt(sum(x * x for x in dd))
    if length == 0:
        return None
    norm = [sv[2], sv[1], sv[0]]  # x, y, z
    dmm = F(f32(step_mm))
    step = [dd[a] / length * dmm / norm[a] for a in range(3)]
    j = F(f32(jitter))
    st_n = [st[a] / norm[a] + Fr(1, 2) + j * step[a] for a in range(3)]
    en_n = [en[a] / norm[a] + Fr(1, 2) + step[a] for a in range(3)]
    clamp = lambda q: min(max(q, Fr(0)), Fr(1))
    return [clamp(q) for q in st_n], step, [clamp(q) for q in en_n]


def exact_count(start, step, end, S: int, window: float = 2e-5):
    """(count, tie) for the voxel break rule; tie=True if a decision lies within ``window``."""
    eps = F(f32(1e-5))
    tie = False
    for i in range(S):
        stop = False
        for a in range(3):
            if step[a] == 0:
                continue
            p = start[a] + i * step[a]
            margin = (p - eps - end[a]) if step[a] > 0 else (end[a] - (p + eps))
            if abs(margin) < window:
                tie = True
            if margin > 0:
                stop = True
        if stop:
            return i, tie
    return S, tie
