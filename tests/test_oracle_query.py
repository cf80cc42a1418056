"""Pins the oracle's export queries (oracle/query.py) with closed forms, torch's CPU operations and exact
arithmetic; checks the product's export extents against it (host logic, no GPU)."""
from fractions import Fraction

import numpy as np
import torch

from oracle import query as oq


def test_linspace_device_closed_forms():
    # endpoints are exact, grids of exactly representable points are reproduced exactly
    for a, b, n in [(0.0, 1.0, 5), (0.0, 1.0, 9), (-1.0, 1.0, 17), (0.25, 0.75, 3), (0.0, 1.0, 2)]:
        got = oq.linspace_device(a, b, n)
        exact = [a + (b - a) * i / (n - 1) for i in range(n)]
        assert got.dtype == np.float32 and got.size == n
        assert np.array_equal(got, np.array(exact, np.float32)), (a, b, n)
    assert np.array_equal(oq.linspace_device(0.3, 0.9, 1), np.array([0.3], np.float32))
    # within an ulp or two of the exact rational grid everywhere
    for a, b, n in [(0.013, 0.911, 37), (0.5 / 256, 1 - 0.5 / 256, 256), (10 / 40, 24 / 40, 8), (0.0, 1.0, 512)]:
        got = oq.linspace_device(a, b, n)
        a32, b32 = Fraction(float(np.float32(a))), Fraction(float(np.float32(b)))
        for i in range(n):
            exact = a32 + (b32 - a32) * Fraction(i, n - 1)
            ulp = float(np.spacing(np.float32(float(exact))))
            assert abs(float(Fraction(float(got[i])) - exact)) <= 2 * ulp, (a, b, n, i)
        assert got[0] == np.float32(a) and got[-1] == np.float32(b)
        # and within two ulps of torch's CPU rule (another rounding of the same grid)
        cpu = oq.linspace_cpu(a, b, n)
        assert np.all(np.abs(got.astype(np.float64) - cpu) <= 2 * np.spacing(np.abs(cpu).astype(np.float32)))


def test_export_extent_known_answers():
    # full grid: z from nVoxel[0], y and x both from nVoxel[1] (static_export.py:53-56)
    assert oq.export_extent((64, 48, 40), binning=2) == ((32, 24, 24), (0, 1), (0, 1), (0, 1))
    # ROIs: [i0, i1) over the grid padded by sample_outside in y/x; z unpadded
    e = oq.export_extent((64, 48, 40), binning=2, ROIx=[8, 24], ROIz=[4, 20])
    assert e == ((8, 24, 8), (4 / 64, 20 / 64), (0, 1), (8 / 40, 24 / 40))
    e = oq.export_extent((64, 48, 40), binning=1, ROIx=[8, 24], ROIy=[6, 30], sample_outside=4)
    assert e == ((64, 24, 16), (0, 1), ((6 - 4) / 56, (30 - 4) / 56), ((8 - 4) / 48, (24 - 4) / 48))


def test_product_export_extent_matches_oracle():
    from nect_amd.export import export_extent

    rng = np.random.default_rng(0)
    for _ in range(200):
        nv = tuple(int(v) for v in rng.integers(8, 300, 3))
        b = int(rng.integers(1, 5))
        rm = int(rng.integers(0, 6))
        roi = [None if rng.random() < 0.4 else sorted(int(v) for v in rng.integers(0, nv[k], 2)) for k in range(3)]
        args = dict(binning=b, ROIx=roi[2], ROIy=roi[1], ROIz=roi[0], sample_outside=rm)
        assert export_extent(nv, **args) == oq.export_extent(nv, **args)


def test_create_volume_axes_are_voxel_centres():
    z, y, x = oq.create_volume_axes((20, 16, 12))
    assert (z.size, y.size, x.size) == (20, 16, 12)
    assert z[0] == np.float32(0.5 / 20) and z[-1] == np.float32(1 - 0.5 / 20)
    # padded by sample_outside and cut back: the same points as the padded grid's inner ones
    z2, y2, x2 = oq.create_volume_axes((20, 16, 12), sample_outside=3)
    assert y2.size == 16 and x2.size == 12
    assert np.array_equal(y2, oq.linspace_device(0.5 / 22, 1 - 0.5 / 22, 22)[3:-3])
    pts = oq.grid_points((z[:2], y[:3], x[:4]), t=0.25)
    assert pts.shape == (24, 4) and np.all(pts[:, 3] == np.float32(0.25))
    assert np.array_equal(pts[5, :3], np.array([z[0], y[1], x[1]], np.float32))


def test_rescale_divide_is_torchs_cpu_arithmetic():
    rng = np.random.default_rng(1)
    mu = (rng.random(10000) * 3).astype(np.float32)
    l_max, span, lo = 181.83, 2.75 - 0.31, 0.31
    ref = ((torch.from_numpy(mu) / l_max) * span + lo).numpy()
    assert np.array_equal(oq.rescale(mu, l_max, span, lo, "divide"), ref)
    # mu rounded to half first (a half-precision network's output, then .float())
    ref16 = ((torch.from_numpy(mu).half().float() / l_max) * span + lo).numpy()
    assert np.array_equal(oq.rescale(mu, l_max, span, lo, "divide", mu_dtype="float16"), ref16)


def test_rescale_reciprocal_within_ulps_and_distinct_from_divide():
    rng = np.random.default_rng(2)
    mu = (rng.random(20000) * 3).astype(np.float32)
    l_max, span, lo = 181.83, 1.0, 0.0
    rec = oq.rescale(mu, l_max, span, lo, "reciprocal")
    div = oq.rescale(mu, l_max, span, lo, "divide")
    exact = mu.astype(np.float64) / np.float64(np.float32(l_max))
    assert np.all(np.abs(rec - exact) <= 2 * np.spacing(np.abs(exact).astype(np.float32)))
    assert np.count_nonzero(rec != div) > 0          # the two rules differ in the last bit on some values
    h = oq.rescale(mu, l_max, 2.0, 0.5, "reciprocal_f16")
    assert np.array_equal(h, h.astype(np.float16).astype(np.float32))


def test_export_timesteps():
    assert oq.export_timesteps([0, 5, 10], 20) == [0.0, 0.25, 0.5]
    assert oq.export_timesteps([0.1, 0.5, 1.0], 20) == [0.1, 0.5, 1.0]
