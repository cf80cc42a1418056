"""CPU checks of the synthetic data generation: union chords through spheres (closed forms and fine
sampling), the packed-sphere phantom's porosity, Poisson noise statistics."""
import math

import numpy as np
import torch

from nect_amd import phantom


def _fine_length(centres, radii, p1, p2, n=200_001):
    t = torch.linspace(0, 1, n, dtype=torch.float64)[:, None]
    pts = p1[None] + t * (p2 - p1)[None]
    inside = ((pts[:, None, :] - centres[None]) ** 2).sum(-1) <= radii[None] ** 2
    frac = inside.any(1).double().mean().item()
    return frac * float((p2 - p1).norm())


def test_union_chords_closed_forms_and_fine_sampling():
    c = torch.tensor([[0.0, 0.0, 0.0]], dtype=torch.float64)
    r = torch.tensor([2.0], dtype=torch.float64)
    P1 = torch.tensor([[-10.0, 0.0, 0.0], [-10.0, 1.0, 0.0], [-10.0, 3.0, 0.0], [0.0, 0.0, 0.0]], dtype=torch.float64)
    P2 = torch.tensor([[10.0, 0.0, 0.0], [10.0, 1.0, 0.0], [10.0, 3.0, 0.0], [10.0, 0.0, 0.0]], dtype=torch.float64)
    got = phantom.union_chords(c, r, P1, P2)
    assert torch.allclose(got, torch.tensor([4.0, 2 * math.sqrt(3), 0.0, 2.0], dtype=torch.float64), atol=1e-12)
    # two copies of one sphere count once; two disjoint spheres add
    c2 = torch.tensor([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [5.0, 0.0, 0.0]], dtype=torch.float64)
    r2 = torch.tensor([2.0, 2.0, 1.0], dtype=torch.float64)
    assert abs(float(phantom.union_chords(c2, r2, P1[:1], P2[:1])) - 6.0) < 1e-12
    # random overlapping spheres against fine sampling of the segment
    rng = np.random.default_rng(0)
    cs = torch.from_numpy(rng.uniform(-4, 4, (12, 3)))
    rs = torch.from_numpy(rng.uniform(0.5, 2.5, 12))
    for k in range(5):
        p1 = torch.from_numpy(rng.uniform(-8, 8, 3)) + torch.tensor([-10.0, 0, 0], dtype=torch.float64)
        p2 = torch.from_numpy(rng.uniform(-8, 8, 3)) + torch.tensor([10.0, 0, 0], dtype=torch.float64)
        got = float(phantom.union_chords(cs, rs, p1[None], p2[None]))
        assert abs(got - _fine_length(cs, rs, p1, p2)) < 2e-3, k


def test_relaxed_packing_reaches_the_target_porosity():
    ph = phantom.relaxed_packed_spheres(128, 0.5, seed=1)
    por = phantom.union_porosity(ph, 128, 0.5)
    assert abs(por - 0.35) < 0.02, por
    # grains inside the pore cylinder
    R_in, _, half_h = ph.cylinder
    rho = ph.sphere_centres[:, :2].norm(dim=1)
    assert torch.all(rho + ph.sphere_radii <= R_in + 1e-9)
    assert torch.all(ph.sphere_centres[:, 2].abs() + ph.sphere_radii <= half_h + 1e-9)


def test_poisson_log_statistics():
    p = torch.full((200_000,), 2.0)
    noisy = phantom.poisson_log(p, 1e4, generator=torch.Generator().manual_seed(0))
    # -log(N / I0) with N ~ Poisson(I0 e^-p): mean ~ p, variance ~ 1 / (I0 e^-p)
    assert abs(float(noisy.mean()) - 2.0) < 2e-3
    assert abs(float(noisy.var()) - 1 / (1e4 * math.exp(-2.0))) / (1 / (1e4 * math.exp(-2.0))) < 0.05


def _mu_points(ph, pts):
    """Attenuation at points [n, 3] (x, y, z mm) of a relaxed-packing / brine-front phantom at one time,
    evaluated point by point (the brute-force counterpart of the chord formulas)."""
    R_in, R_out, half_h = ph.cylinder
    rho2 = pts[:, 0] ** 2 + pts[:, 1] ** 2
    inz = pts[:, 2].abs() <= half_h
    wall, pore = float(ph.mus[0]), float(ph.mus[2])
    mu = torch.where(inz & (rho2 <= R_out ** 2) & (rho2 > R_in ** 2), wall, 0.0)
    in_pore = inz & (rho2 <= R_in ** 2)
    d2 = ((pts[:, None, :] - ph.sphere_centres[None]) ** 2).sum(-1)
    grain = (d2 <= ph.sphere_radii[None] ** 2).any(1)
    mu = mu + torch.where(in_pore, pore, 0.0) + torch.where(grain, ph.grain_excess, 0.0)
    h = getattr(ph, "brine_level", None)
    if h is not None:
        mu = mu + torch.where(in_pore & ~grain & (pts[:, 2] < h), ph.brine_excess, 0.0)
    return mu


def _riemann(ph, P1, P2, n=20000):
    s = (torch.arange(n, dtype=torch.float64) + 0.5) / n
    out = []
    for a, b in zip(P1, P2):
        pts = a[None] + s[:, None] * (b - a)[None]
        out.append(float(_mu_points(ph, pts).sum()) * float((b - a).norm()) / n)
    return torch.tensor(out, dtype=torch.float64)


def test_brine_front_projection_matches_point_sampling():
    """Cone-beam line integrals of the brine-front phantom (shrinking soluble grains, brine below the front,
    union of grains) equal a fine Riemann sum of the attenuation sampled point by point along the same rays,
    at three times; at t = 0 the phantom is the static packing (no brine, full grains)."""
    from nect_amd import phantom

    base = phantom.relaxed_packed_spheres(32, 1.0, seed=3, iterations=60)
    ph = phantom.brine_front(base, 32, 1.0, soluble_fraction=0.5, seed=1)
    pix = torch.tensor([16 * 32 + 16, 10 * 32 + 7, 20 * 32 + 25, 5 * 32 + 16, 27 * 32 + 3, 16 * 32 + 30])
    angle = 0.6
    P1, P2 = phantom.ray_endpoints("cone", (32, 32), (1.4, 1.4), 60.0, 120.0, angle, pix)
    for t in (0.0, 0.45, 0.8):
        got = phantom.project(ph, "cone", (32, 32), (1.4, 1.4), 60.0, 120.0, [angle], device="cpu",
                              times=[t])[0].double().flatten()[pix]
        ref = _riemann(phantom.at_time(ph, t), P1, P2)
        assert torch.allclose(got, ref, rtol=2e-3, atol=2e-3), (t, got, ref)
    static = phantom.project(base, "cone", (32, 32), (1.4, 1.4), 60.0, 120.0, [angle], device="cpu")
    t0 = phantom.project(ph, "cone", (32, 32), (1.4, 1.4), 60.0, 120.0, [angle], device="cpu", times=[0.0])
    assert torch.allclose(static, t0, rtol=1e-6, atol=1e-6)
    # soluble grains are gone once the front has passed them; the others keep their radius
    late = phantom.at_time(ph, 1.0)
    assert torch.all(late.sphere_radii[ph.soluble] == 0)
    assert torch.equal(late.sphere_radii[~ph.soluble], base.sphere_radii[~ph.soluble])


def test_union_chords_by_row_equals_union_chords():
    from nect_amd import phantom

    base = phantom.relaxed_packed_spheres(32, 1.0, seed=5, iterations=30)
    pix = torch.arange(24 * 24)
    P1, P2 = phantom.ray_endpoints("cone", (24, 24), (1.7, 1.7), 60.0, 120.0, 1.1, pix)
    a = phantom.union_chords(base.sphere_centres, base.sphere_radii, P1, P2)
    b = phantom.union_chords_by_row(base.sphere_centres, base.sphere_radii, P1, P2, 24, rows_per_block=3)
    assert torch.allclose(a, b, rtol=1e-12, atol=1e-12)
    c = phantom.union_chords(base.sphere_centres, base.sphere_radii, P1, P2, z_below=2.5)
    d = phantom.union_chords_by_row(base.sphere_centres, base.sphere_radii, P1, P2, 24, z_below=2.5)
    assert torch.allclose(c, d, rtol=1e-12, atol=1e-12)
    assert torch.all(c <= a + 1e-12)


def test_ellipse_slab_projection_matches_voxel_and_point_sums():
    """Parallel-beam line integrals of the ellipse slab (elliptic cylinders with their own z-extent) equal a
    fine point sum along the rays; about per_slice ellipses cross each slice."""
    from nect_amd import phantom

    ph = phantom.ellipse_slab(64, 16, 1.0, per_slice=30, seed=2)
    zr = ph.z_ranges
    zs = torch.arange(16, dtype=torch.float64) - 7.5
    crossing = ((zs[:, None] >= zr[None, :, 0]) & (zs[:, None] <= zr[None, :, 1])).sum(1).double()
    assert 15 < float(crossing.mean()) < 45
    pix = torch.tensor([0 * 64 + 3, 5 * 64 + 31, 9 * 64 + 50, 15 * 64 + 12, 7 * 64 + 63])
    P1, P2 = phantom.ray_endpoints("parallel", (16, 64), (1.0, 1.0), 0.0, 0.0, 0.3, pix, reach=80.0)
    got = phantom.line_integrals(ph, P1, P2)
    n = 40000
    s = (torch.arange(n, dtype=torch.float64) + 0.5) / n
    ref = []
    for a, b in zip(P1, P2):
        pts = a[None] + s[:, None] * (b - a)[None]
        q = torch.einsum("nd,kde->nke", pts, ph.rots) - torch.einsum("kd,kde->ke", ph.centres, ph.rots)[None]
        inside = ((q / ph.axes[None]) ** 2).sum(-1) <= 1
        inside &= (pts[:, 2:3] >= zr[None, :, 0]) & (pts[:, 2:3] <= zr[None, :, 1])
        ref.append(float((inside.double() * ph.mus[None]).sum()) * float((b - a).norm()) / n)
    assert torch.allclose(got, torch.tensor(ref, dtype=torch.float64), rtol=2e-3, atol=2e-3)
