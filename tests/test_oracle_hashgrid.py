"""Pins the oracle hash grid (oracle/hashgrid.py) with closed forms and brute force."""
import math

import numpy as np
import pytest

from oracle.hashgrid import PRIMES, GridSpec, corner_index, corners, encode, backward, level_position


def spec_cfg2():
    return GridSpec(3, 16, 2, 19, 16, 1.5)


def test_level_table_closed_form():
    s = spec_cfg2()
    # 16 * 1.5^l - 1 is exact in binary for these levels
    assert list(s.scale[:6]) == [15.0, 23.0, 35.0, 53.0, 80.0, 120.5]
    assert list(s.resolution[:6]) == [16, 24, 36, 54, 81, 122]
    T = 1 << 19
    assert list(s.level_size[:4]) == [16 ** 3, 24 ** 3, 36 ** 3, 54 ** 3]   # dense, multiples of 8
    assert all(int(v) == T for v in s.level_size[4:])                        # 81^3 > 2^19
    assert s.offset[0] == 0 and s.offset[-1] == sum(int(v) for v in s.level_size)
    # finest resolution ~ 16 * 1.5^15
    assert abs(s.resolution[15] - (math.ceil(16 * 1.5 ** 15 - 1) + 1)) == 0
    assert s.n_params == s.offset[-1] * 2
    assert abs(s.n_params / 2 - 6.5e6) < 0.5e6   # about 13.0M parameters (F=2)


def test_level_table_4d_and_odd_sizes():
    s = GridSpec(4, 16, 2, 21, 16, 1.5)
    assert s.level_size[0] == 16 ** 4 and s.level_size[1] == 24 ** 4 and s.level_size[2] == 36 ** 4
    assert s.level_size[3] == 1 << 21
    s2 = GridSpec(3, 4, 2, 12, 5, 1.3)
    for l in range(4):
        d = int(s2.resolution[l]) ** 3
        assert s2.level_size[l] == min(((d + 7) // 8) * 8, 1 << 12)
        assert s2.level_size[l] % 8 == 0 or s2.level_size[l] == 1 << 12


def test_dense_index_brute_force():
    s = GridSpec(3, 2, 2, 16, 6, 1.5)   # res 6, 9: dense
    rng = np.random.default_rng(0)
    for l in range(2):
        res = int(s.resolution[l])
        c = rng.integers(0, res, size=(200, 3)).astype(np.uint64)
        got = corner_index(s, l, c)
        for i in range(200):
            expect = (int(c[i, 0]) + int(c[i, 1]) * res + int(c[i, 2]) * res * res) % int(s.level_size[l])
            assert int(got[i]) == expect


def test_hash_index_exact_integers():
    s = spec_cfg2()
    l = 10
    rng = np.random.default_rng(1)
    c = rng.integers(0, int(s.resolution[l]) + 1, size=(300, 3)).astype(np.uint64)
    got = corner_index(s, l, c)
    for i in range(300):
        h = 0
        for d in range(3):
            h ^= (int(c[i, d]) * PRIMES[d]) % (1 << 32)
        assert int(got[i]) == h % int(s.level_size[l])


def test_hash_index_4d_exact_integers():
    s = GridSpec(4, 16, 2, 21, 16, 1.5)
    l = 12
    rng = np.random.default_rng(2)
    c = rng.integers(0, int(s.resolution[l]) + 1, size=(300, 4)).astype(np.uint64)
    got = corner_index(s, l, c)
    for i in range(300):
        h = 0
        for d in range(4):
            h ^= (int(c[i, d]) * PRIMES[d]) % (1 << 32)
        assert int(got[i]) == h % (1 << 21)


def test_partition_of_unity():
    s = spec_cfg2()
    x = np.random.default_rng(3).random((500, 3), dtype=np.float32)
    _, w = corners(s, x)
    assert np.all(w >= 0)
    assert np.allclose(w.astype(np.float64).sum(-1), 1.0, atol=2e-6)


def test_origin_has_two_half_corners_per_dim():
    s = spec_cfg2()
    x = np.zeros((1, 3), np.float32)
    c, f = level_position(s, 0, x)
    assert np.all(c == 0) and np.all(f == np.float32(0.5))
    _, w = corners(s, x)
    assert np.allclose(w, 0.125)


def test_affine_table_known_answer():
    """Multilinear interpolation reproduces an affine function of the grid coordinate."""
    s = GridSpec(3, 3, 2, 20, 8, 1.5)   # dense levels
    table = np.zeros(s.n_params)
    coef = np.array([[0.3, -1.1, 2.0, 0.7], [1.5, 0.25, -0.5, 3.0]])
    for l in range(3):
        res = int(s.resolution[l])
        n = int(s.level_size[l])
        e = np.arange(n)
        g0, g1, g2 = e % res, (e // res) % res, e // (res * res)
        for f in range(2):
            a, b0, b1, b2 = coef[f]
            table[(s.offset[l] + e) * 2 + f] = a + b0 * g0 + b1 * g1 + b2 * g2
    rng = np.random.default_rng(4)
    x = (rng.random((300, 3)) * 0.8).astype(np.float32)   # keep corner+1 < res: no wrap
    out = encode(s, table, x)
    for l in range(3):
        c, f = level_position(s, l, x)
        p = c.astype(np.float64) + f.astype(np.float64)
        for ff in range(2):
            a, b0, b1, b2 = coef[ff]
            expect = a + b0 * p[:, 0] + b1 * p[:, 1] + b2 * p[:, 2]
            assert np.allclose(out[:, l * 2 + ff], expect, rtol=0, atol=1e-5 * (1 + np.abs(expect)))


def test_backward_is_adjoint_of_encode():
    """<dout, encode(table)> = <backward(dout), table> for random table and dout."""
    s = GridSpec(3, 6, 2, 12, 4, 1.7)
    rng = np.random.default_rng(5)
    x = rng.random((200, 3), dtype=np.float32)
    table = rng.standard_normal(s.n_params)
    dout = rng.standard_normal((200, 12))
    lhs = np.sum(dout * encode(s, table, x))
    rhs = np.dot(backward(s, x, dout), table)
    assert abs(lhs - rhs) < 1e-9 * (1 + abs(lhs))


@pytest.mark.parametrize("l", [0, 3, 9])
def test_finite_difference_of_features_in_x(l):
    """Inside a cell the feature is multilinear in p: FD slope equals the corner difference formula."""
    s = GridSpec(3, 10, 2, 14, 4, 1.5)
    rng = np.random.default_rng(6)
    table = rng.standard_normal(s.n_params)
    x = np.array([[0.31, 0.52, 0.77]], np.float32)
    base = encode(s, table, x)[0, l * 2]
    c, f = level_position(s, l, x)
    # move inside the cell along dim 0 by a representable step
    h = np.float32(1e-3) / s.scale[l]
    x2 = x.copy()
    x2[0, 0] += h
    c2, f2 = level_position(s, l, x2)
    if np.any(c2 != c):
        pytest.skip("step crossed a cell")
    d = encode(s, table, x2)[0, l * 2] - base
    # slope along p0 from the 4 edge pairs
    idx, w = corners(s, x)
    tab = table.reshape(-1, 2)
    slope = 0.0
    for k in range(8):
        if k & 1:
            continue
        w_rest = 1.0
        for dd in (1, 2):
            w_rest *= f[0, dd] if (k >> dd) & 1 else 1 - f[0, dd]
        v0 = tab[s.offset[l] + idx[0, l, k], 0]
        v1 = tab[s.offset[l] + idx[0, l, k | 1], 0]
        slope += w_rest * (v1 - v0)
    dp = float(f2[0, 0]) - float(f[0, 0])
    assert abs(d - slope * dp) < 1e-5 * (1 + abs(d))


def _round_f32_exact(q):
    """Round-to-nearest-even of a Fraction to float32, by comparing exact distances to the neighbours."""
    from fractions import Fraction

    c = np.float32(float(q))
    cands = [np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))]
    best = None
    for v in cands:
        d = abs(Fraction(float(v)) - q)
        key = (d, int(np.array(v, np.float32).view(np.uint32)) & 1)   # ties: even significand first
        if best is None or key < best[0]:
            best = (key, v)
    return np.float32(best[1])


def test_fma32_matches_exact_rationals():
    """fma32 (one rounding of a*b + c) equals exact rational arithmetic rounded once: random operands, the
    position operands of every level, and constructed double-rounding ties (the float64 sum lands exactly
    half-way between two float32 values while the true value does not)."""
    from fractions import Fraction

    from oracle.hashgrid import fma32

    rng = np.random.default_rng(3)
    a = rng.random(400, dtype=np.float32) * np.float32(8000)
    b = rng.random(400, dtype=np.float32)
    c = np.full(400, np.float32(0.5))
    # double-rounding ties: a*b = 3*2^-25 * (1 + 2^-20)(1 - 2^-20) = 3*2^-25 - 3*2^-65; + 0.5 rounds in float64
    # to 0.5 + 3*2^-25, the midpoint of 0.5 + 2^-24 and 0.5 + 2^-23, while the true value is below it
    ta = np.array([np.float32(3 * 2.0 ** -25 * (1 + 2.0 ** -20)), np.float32(2.0 ** -25 * (1 + 2.0 ** -20))],
                  np.float32)
    tb = np.array([np.float32(1 - 2.0 ** -20), np.float32(1 - 2.0 ** -20)], np.float32)
    a = np.concatenate([a, ta])
    b = np.concatenate([b, tb])
    c = np.concatenate([c, np.full(2, np.float32(0.5))])
    got = fma32(a, b, c)
    for i in range(a.size):
        q = Fraction(float(a[i])) * Fraction(float(b[i])) + Fraction(float(c[i]))
        assert got[i] == _round_f32_exact(q), (i, a[i], b[i])
    # the tie case really is one: two roundings (through float64) give the other neighbour
    naive = (ta.astype(np.float64) * tb.astype(np.float64) + 0.5).astype(np.float32)
    assert naive[0] != got[-2]


def test_position_rule_single_rounding_where_rules_disagree():
    """The oracle's position is fma32's (exact rational check), next to cell boundaries and on random points.
    Rounding the product before adding 0.5 (two roundings) gives another fraction -- so other corner weights
    -- on a fraction of the coordinates; the cell is the same under both rules (rounding is monotone and
    k - 0.5, k are representable), so indices alone cannot tell the rules apart, the weights can."""
    from fractions import Fraction

    from tests.helpers import rounding_edge_points

    rng = np.random.default_rng(5)
    for spec in (spec_cfg2(), GridSpec(4, 16, 2, 21, 16, 1.5)):
        x = np.concatenate([rounding_edge_points(spec, 48, seed=1), rng.random((4000, spec.n_dims), dtype=np.float32)])
        frac_disagree = 0
        for l in range(spec.n_levels):
            s = np.float32(spec.scale[l])
            c, f = level_position(spec, l, x)
            two = ((s * x).astype(np.float32) + np.float32(0.5)).astype(np.float32)
            assert np.array_equal(np.floor(two).astype(np.uint64), c)
            frac_disagree += int(np.count_nonzero((two - np.floor(two)).astype(np.float32) != f))
            rows = np.concatenate([np.arange(l * 48, l * 48 + 8), 16 * 48 + rng.integers(0, 4000, 8)])
            for i in rows:
                for d in range(spec.n_dims):
                    p = _round_f32_exact(Fraction(float(s)) * Fraction(float(x[i, d])) + Fraction(1, 2))
                    assert int(c[i, d]) == int(np.floor(p)) and f[i, d] == np.float32(p - np.floor(p))
        assert frac_disagree > 0
