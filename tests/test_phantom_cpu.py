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
