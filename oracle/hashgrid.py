"""ORACLE (test infrastructure): multiresolution hash-grid encoding in numpy.

The reference reaches the encoding through tiny-cuda-nn's ``HashGrid``
(nect/network/networks.py:28-65, configured by nect/config.py:86-112).
tiny-cuda-nn is an un-vendored, unpinned third-party dependency (installed from
``git+https://github.com/NVlabs/tiny-cuda-nn/#subdirectory=bindings/torch``,
README.md:66,75,85), absent here.  This module restates its published
algorithm (Instant-NGP, Mueller et al. 2022: per-level scale, dense levels while
res^D fits, otherwise the xor of coordinate*prime, modulo the level size;
(multi)linear interpolation of the 2^D corners) as the build's specification,
with the rounding rules fixed explicitly:

  scale_l  = float32(base * pls**l - 1)   (float64, rounded once), where
             pls = float32(per_level_scale), the width tiny-cuda-nn stores it in
  res_l    = ceil(scale_l) + 1
  size_l   = min(next_multiple_of_8(res_l**D), 2**log2_hashmap_size)
  p_d      = fma32(scale_l, x_d, 0.5)  (one rounding, tiny-cuda-nn's
             fmaf(scale, x, 0.5f)),  c_d = floor(p_d),  f_d = p_d - c_d
  weight_k = prod_{d=0..D-1} (f_d if bit d of k else 1 - f_d)   (float32, in order)
  index_k  = sum_d c'_d res^d (dense: while the stride <= size), else
             xor_d c'_d * prime_d mod 2^32; then % size_l

Parity pins (tests/test_oracle_hashgrid.py): the level table against closed
forms; dense indices by brute force; hash values by exact Python integers;
partition of unity; an affine-table known answer (multilinear interpolation
reproduces affine functions exactly); x = 0 -> two corners weighted 1/2;
fma32 against exact rational arithmetic, including inputs where rounding the
product first (two roundings) picks another cell.
Parity with tiny-cuda-nn itself (its exp2f-derived scales, half-precision
accumulation) is unpinned: the library is not importable here.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

PRIMES = (1, 2654435761, 805459861, 3674653429)
MASK32 = np.uint64(0xFFFFFFFF)


@dataclass
class GridSpec:
    n_dims: int
    n_levels: int
    n_features: int
    log2_hashmap_size: int
    base_resolution: int
    per_level_scale: float

    def __post_init__(self):
        T = 1 << self.log2_hashmap_size
        self.scale = np.zeros(self.n_levels, np.float32)
        self.resolution = np.zeros(self.n_levels, np.int64)
        self.level_size = np.zeros(self.n_levels, np.int64)
        self.offset = np.zeros(self.n_levels + 1, np.int64)
        off = 0
        pls = float(np.float32(self.per_level_scale))
        for l in range(self.n_levels):
            sc = np.float32(self.base_resolution * math.pow(pls, l) - 1.0)
            res = int(math.ceil(float(sc))) + 1
            dense = res ** self.n_dims
            size = min(((dense + 7) // 8) * 8, T)
            self.scale[l] = sc
            self.resolution[l] = res
            self.level_size[l] = size
            self.offset[l] = off
            off += size
        self.offset[self.n_levels] = off
        self.n_params = off * self.n_features


def fma32(a, b, c) -> np.ndarray:
    """float32 a*b + c with a single rounding (IEEE fused multiply-add, round to nearest even), exactly.

    The product of two float32 values is exact in float64 (48 significant bits); the float64 sum s may round,
    and TwoSum gives its exact error e (s + e = a*b + c).  Rounding s to float32 is then the correct rounding
    of a*b + c unless s lies exactly half-way between two float32 values while e != 0 (a double-rounding
    tie): then the true value lies on the side of e."""
    a64 = np.asarray(a, dtype=np.float32).astype(np.float64)
    b64 = np.asarray(b, dtype=np.float32).astype(np.float64)
    c64 = np.asarray(c, dtype=np.float32).astype(np.float64)
    p = a64 * b64
    s = p + c64
    bb = s - p
    e = (p - (s - bb)) + (c64 - bb)
    r = s.astype(np.float32)
    r64 = r.astype(np.float64)
    # the float32 neighbour of r on the side of s, and the midpoint between them
    other = np.nextafter(r, np.where(s > r64, np.float32(np.inf), np.float32(-np.inf)).astype(np.float32))
    mid = (r64 + other.astype(np.float64)) * 0.5
    tie = (s != r64) & (s == mid) & (e != 0)
    if np.any(tie):
        lo = np.minimum(r, other)
        hi = np.maximum(r, other)
        r = np.where(tie, np.where(e > 0, hi, lo), r).astype(np.float32)
    return np.asarray(r, dtype=np.float32)


def level_position(spec: GridSpec, l: int, x: np.ndarray):
    """(c uint64 [n, D], f float32 [n, D]): p = fma32(scale, x, 0.5), c = floor(p), f = p - c."""
    x = np.asarray(x, dtype=np.float32)
    p = fma32(np.float32(spec.scale[l]), x, np.float32(0.5))
    fl = np.floor(p)
    return fl.astype(np.int64).astype(np.uint64), (p - fl).astype(np.float32)


def corner_index(spec: GridSpec, l: int, c: np.ndarray) -> np.ndarray:
    """Index within level l of integer corners c [n, D] (uint64)."""
    res = np.uint64(spec.resolution[l])
    size = int(spec.level_size[l])
    D = spec.n_dims
    stride = 1
    index = np.zeros(c.shape[0], np.uint64)
    dense = True
    for d in range(D):
        if stride > size:
            dense = False
            break
        index = (index + c[:, d] * np.uint64(stride)) & MASK32
        stride *= int(res)
    if stride > size:
        dense = False
    if not dense:
        h = np.zeros(c.shape[0], np.uint64)
        for d in range(D):
            h ^= (c[:, d] * np.uint64(PRIMES[d])) & MASK32
        index = h
    return (index % np.uint64(size)).astype(np.uint32)


def corners(spec: GridSpec, x: np.ndarray):
    """index uint32 [n, L, 2^D] (within level) and weight float32 [n, L, 2^D]."""
    x = np.asarray(x, dtype=np.float32)
    n, D = x.shape
    K = 1 << D
    idx = np.zeros((n, spec.n_levels, K), np.uint32)
    wts = np.zeros((n, spec.n_levels, K), np.float32)
    for l in range(spec.n_levels):
        c, f = level_position(spec, l, x)
        for k in range(K):
            w = np.ones(n, np.float32)
            cc = c.copy()
            for d in range(D):
                if (k >> d) & 1:
                    w = (w * f[:, d]).astype(np.float32)
                    cc[:, d] = cc[:, d] + np.uint64(1)
                else:
                    w = (w * (np.float32(1.0) - f[:, d]).astype(np.float32)).astype(np.float32)
            idx[:, l, k] = corner_index(spec, l, cc)
            wts[:, l, k] = w
    return idx, wts


def encode(spec: GridSpec, table: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Features [n, L*F] in float64 from table [n_params] (any float dtype)."""
    idx, wts = corners(spec, x)
    F = spec.n_features
    tab = np.asarray(table, dtype=np.float64).reshape(-1, F)
    n = idx.shape[0]
    out = np.zeros((n, spec.n_levels * F), np.float64)
    for l in range(spec.n_levels):
        rows = tab[spec.offset[l] + idx[:, l, :].astype(np.int64)]          # [n, K, F]
        out[:, l * F:(l + 1) * F] = np.einsum("nk,nkf->nf", wts[:, l, :].astype(np.float64), rows)
    return out


def backward(spec: GridSpec, x: np.ndarray, dout: np.ndarray) -> np.ndarray:
    """Gradient of sum(dout * encode) w.r.t. the table, float64 [n_params]."""
    idx, wts = corners(spec, x)
    F = spec.n_features
    grad = np.zeros((spec.offset[-1], F), np.float64)
    dout = np.asarray(dout, dtype=np.float64)
    for l in range(spec.n_levels):
        rows = spec.offset[l] + idx[:, l, :].astype(np.int64)               # [n, K]
        contrib = wts[:, l, :, None].astype(np.float64) * dout[:, None, l * F:(l + 1) * F]
        np.add.at(grad, rows.reshape(-1), contrib.reshape(-1, F))
    return grad.reshape(-1)
