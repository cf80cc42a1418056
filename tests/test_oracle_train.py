"""Pins the trainer oracle's line integral against analytic chords, and its gradient against FD."""
import numpy as np
import torch

from oracle import analytic as an
from oracle import hashgrid as hg
from oracle.geometry import OracleGeometry
from oracle.sampler import oracle_sample
from oracle.train import integrate, train_step


def test_line_integral_of_exact_model_matches_analytic():
    g = OracleGeometry(mode="cone", nDetector=(32, 32), dDetector=(1.0, 1.0), DSD=400.0, DSO=200.0,
                       nVoxel=(32, 32, 32), dVoxel=(0.5, 0.5, 0.5))
    f = g.c_struct_fields()
    S = 256
    step = g.max_distance_traveled / S
    centres = np.array([[0.5, -1.0, 0.3]])
    axes = np.array([[4.0, 5.0, 3.0]])
    rots = [np.eye(3)]
    mus = [0.8]
    B = 1024
    rng = np.random.default_rng(0)
    idx = rng.permutation(1024).astype(np.int64)
    for angle in (0.0, 1.1, 4.0):
        pts, dist, cnt = oracle_sample(f, angle, idx, B, 0, S, step, True, np.zeros(B, np.float32),
                                       np.zeros((B, 2), np.float32))
        s_vox = np.array(g.sVoxel)  # z, y, x
        mm = (pts.astype(np.float64) - 0.5) * s_vox          # (z, y, x) mm
        xyz = mm[..., ::-1].reshape(-1, 3)
        mu = an.voxel_value(xyz, centres, axes, rots, mus).reshape(B, S)
        mu[pts.sum(-1) == 0] = 0
        y = integrate(torch.from_numpy(mu), dist, g.max_distance_traveled).numpy() * g.max_distance_traveled
        P1, P2 = an.ray_points("cone", (32, 32), (1.0, 1.0), 200.0, 400.0, angle, idx[:B])
        ref = an.segment_integral(P1, P2, centres, axes, rots, mus)
        # two boundary crossings, each off by at most one spacing
        assert np.all(np.abs(y - ref) <= 2 * step * 0.8 + 1e-9), np.abs(y - ref).max()
        assert ref.max() > 5


def test_train_step_gradient_matches_finite_differences():
    g = OracleGeometry(mode="cone", nDetector=(8, 8), dDetector=(1.0, 1.0), DSD=100.0, DSO=50.0, nVoxel=(8, 8, 8),
                       dVoxel=(0.5, 0.5, 0.5))
    f = g.c_struct_fields()
    spec = hg.GridSpec(3, 4, 2, 10, 4, 1.5)
    mlp_cfg = {"n_params": 64 * 16 + 16 * 64, "n_input_padded": 16, "n_hidden_layers": 1, "n_output_padded": 16,
               "activation": "Tanh", "output_activation": "Softplus"}
    rng = np.random.default_rng(1)
    params = np.concatenate([rng.standard_normal(mlp_cfg["n_params"]) * 0.2,
                             rng.standard_normal(spec.n_params) * 0.1])
    B, S = 16, 12
    idx = rng.permutation(64).astype(np.int64)
    jit = (rng.random(B) - 0.5).astype(np.float32)
    off = np.zeros((B, 2), np.float32)
    tgt = rng.random(B)
    args = (f, 0.4, idx, 0, B, S, g.max_distance_traveled / S, True, jit, off, tgt, g.max_distance_traveled, spec,
            mlp_cfg)
    out = train_step(*args, params=params, loss="L2")
    for k in list(rng.choice(mlp_cfg["n_params"], 6, replace=False)) + list(
            mlp_cfg["n_params"] + np.nonzero(out["grad"][mlp_cfg["n_params"]:])[0][:6]):
        h = 1e-6
        pp, pm = params.copy(), params.copy()
        pp[k] += h
        pm[k] -= h
        fd = (train_step(*args, params=pp)["loss"] - train_step(*args, params=pm)["loss"]) / (2 * h)
        assert abs(fd - out["grad"][k]) < 1e-6 + 1e-4 * abs(fd), (k, fd, out["grad"][k])


def test_oracle_gradient_matches_central_differences_parallel_slab():
    """The oracle's autograd gradient of the whole step against float64 central differences (h = 1e-3) on
    every MLP weight and 256 random touched hash entries, parallel-beam slab geometry."""
    from tests.helpers import FdCase

    c = FdCase(seed=0)
    ref = train_step(c.fields, c.angle, c.idx, 0, c.B, c.S, c.step_mm, True, c.jit, c.off, c.target, c.l_max,
                     c.spec, c.cfg, c.params)
    assert abs(ref["loss"] - c.loss()) < 1e-12
    which = c.fd_indices()
    fd = c.central_differences(which)
    g = ref["grad"][which]
    n_mlp = c.cfg["n_params"]
    e_mlp = np.linalg.norm(fd[:n_mlp] - g[:n_mlp]) / np.linalg.norm(g[:n_mlp])
    e_grid = np.linalg.norm(fd[n_mlp:] - g[n_mlp:]) / np.linalg.norm(g[n_mlp:])
    assert e_mlp < 3e-3, e_mlp
    assert e_grid < 2e-2, e_grid


def test_accumulated_step_reduces_to_single_step():
    """K = 1 is the plain step; K copies of one sub-angle give the same prediction and gradient."""
    from oracle.train import train_step_accumulated

    g = OracleGeometry(mode="cone", nDetector=(8, 8), dDetector=(1.0, 1.0), DSD=100.0, DSO=50.0, nVoxel=(8, 8, 8),
                       dVoxel=(0.5, 0.5, 0.5))
    f = g.c_struct_fields()
    spec = hg.GridSpec(3, 4, 2, 10, 4, 1.5)
    cfg = {"n_params": 64 * 16 + 16 * 64, "n_input_padded": 16, "n_hidden_layers": 1, "n_output_padded": 16,
           "activation": "ReLU", "output_activation": "Softplus"}
    rng = np.random.default_rng(3)
    params = np.concatenate([rng.normal(0, 0.2, cfg["n_params"]), rng.uniform(-0.5, 0.5, spec.n_params)])
    B, S = 16, 16
    idx = rng.permutation(64).astype(np.int64)
    jit = (rng.random(B, dtype=np.float32) - np.float32(0.5))
    off = np.zeros((B, 2), np.float32)
    tgt = rng.random(B)
    step = g.max_distance_traveled / S
    one = train_step(f, 0.4, idx, 0, B, S, step, True, jit, off, tgt, g.max_distance_traveled, spec, cfg, params)
    acc1 = train_step_accumulated(f, [0.4], idx, 0, B, S, step, True, [jit], [off], tgt, g.max_distance_traveled,
                                  spec, cfg, params)
    assert np.array_equal(one["y_pred"], acc1["y_pred"]) and np.array_equal(one["grad"], acc1["grad"])
    acc3 = train_step_accumulated(f, [0.4] * 4, idx, 0, B, S, step, True, [jit] * 4, [off] * 4, tgt,
                                  g.max_distance_traveled, spec, cfg, params)
    assert np.allclose(acc3["y_pred"], one["y_pred"], rtol=1e-13, atol=0)
    assert np.allclose(acc3["grad"], one["grad"], rtol=1e-12, atol=1e-15)
