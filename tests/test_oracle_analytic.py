"""Pins the analytic line-integral oracle and the trainer oracle's line integral."""
import numpy as np

from oracle import analytic as an


def rot(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    Rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    Ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    Rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return Rz @ Ry @ Rx


def test_chord_through_sphere_centre_is_diameter():
    c = np.array([[1.0, -2.0, 0.5]])
    P1 = np.array([[-50.0, -2.0, 0.5], [1.0, -2.0, -40.0]])
    P2 = np.array([[50.0, -2.0, 0.5], [1.0, -2.0, 60.0]])
    got = an.segment_integral(P1, P2, c, np.array([[3.0, 3.0, 3.0]]), [np.eye(3)], [0.7])
    assert np.allclose(got, 0.7 * 6.0, rtol=1e-12)


def test_random_rays_match_numerical_quadrature():
    rng = np.random.default_rng(0)
    K = 4
    centres = rng.uniform(-5, 5, (K, 3))
    axes = rng.uniform(1, 6, (K, 3))
    rots = [rot(*rng.uniform(0, np.pi, 3)) for _ in range(K)]
    mus = rng.uniform(0, 1, K)
    P1 = rng.uniform(-20, 20, (50, 3))
    P2 = rng.uniform(-20, 20, (50, 3))
    got = an.segment_integral(P1, P2, centres, axes, rots, mus)
    s = (np.arange(200000) + 0.5) / 200000
    for i in range(50):
        pts = P1[i] + s[:, None] * (P2[i] - P1[i])
        num = an.voxel_value(pts, centres, axes, rots, mus).mean() * np.linalg.norm(P2[i] - P1[i])
        assert abs(num - got[i]) < 2e-3 * (1 + abs(got[i])), (i, num, got[i])


def test_sphere_projection_is_rotation_invariant():
    c = np.zeros((1, 3))
    a = np.array([[4.0, 4.0, 4.0]])
    pix = np.arange(15 * 15)
    vals = []
    for ang in (0.0, 0.7, 2.1):
        P1, P2 = an.ray_points("cone", (15, 15), (1.0, 1.0), 100.0, 200.0, ang, pix)
        vals.append(an.segment_integral(P1, P2, c, a, [np.eye(3)], [1.0]))
    assert np.allclose(vals[0], vals[1]) and np.allclose(vals[0], vals[2])
    assert vals[0].max() > 7.9
