"""GPU parity of the export queries against the oracle (oracle/query.py): torch's device linspace and the
rescale arithmetic pinned against torch on the device, the query kernel's coordinates and rescale bit-exact,
its values against the float64 model, the export entry points end to end, and alone-vs-mixed serving at the
512^2 / 256^3 / 512^3 request sizes."""
import math

import numpy as np
import pytest
import torch

from oracle import hashgrid as hg
from oracle import query as oq
from oracle.train import model_forward

pytestmark = pytest.mark.gpu

L_MAX, SPAN, LO = 181.83, 2.44, 0.31
MODES = ["divide", "reciprocal", "reciprocal_f16"]


def model(device, dtype, dims=3, hidden=2, log2T=17, seed=0):
    from nect_amd.fused import init_params
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc

    grid = GridDesc(dims, 16, 2, log2T, 16, 1.5)
    mlp = MlpDesc(32, 1, 64, hidden, "LeakyReLU", "Softplus")
    p = init_params(grid, mlp, seed=seed, device=device)
    g = torch.Generator().manual_seed(seed + 7)
    p[mlp.n_params:] = ((torch.rand(grid.n_params, generator=g) * 2 - 1) * 0.3).to(device)
    sh = p.to(dtype)
    spec = hg.GridSpec(dims, 16, 2, log2T, 16, 1.5)
    cfg = {"n_params": mlp.n_params, "n_input_padded": 32, "n_hidden_layers": hidden, "n_output_padded": 16,
           "activation": "LeakyReLU", "output_activation": "Softplus"}
    return grid, mlp, sh[: mlp.n_params], sh[mlp.n_params:], spec, cfg


def module_mu(grid, mlp, w, tab, pts: np.ndarray, device) -> np.ndarray:
    """The network output on points through nect_grid_encode + nect_mlp_forward (float32)."""
    from nect_amd import grid as G
    from nect_amd import mlp as M

    x = torch.from_numpy(pts).to(device)
    return M.forward(mlp, w, G.encode(grid, tab, x))[:, 0].cpu().numpy()


def test_linspace_rule_matches_torch_on_the_device(device):
    print("[query] linspace rule vs torch on the device", flush=True)
    rng = np.random.default_rng(0)
    cases = [(0.0, 1.0, n) for n in (1, 2, 3, 5, 37, 256, 511, 512, 1000)]
    cases += [(0.013, 0.911, 37), (0.5 / 256, 1 - 0.5 / 256, 256), (8 / 40, 24 / 40, 8), (0.25, 0.75, 129),
              (0.1, 0.1, 7), (-0.3, 1.7, 300)]
    cases += [(float(a), float(b), int(n)) for a, b, n in zip(rng.random(40), rng.random(40) + 0.5,
                                                                rng.integers(2, 2000, 40))]
    for a, b, n in cases:
        ref = torch.linspace(a, b, steps=n, device=device).cpu().numpy()
        assert np.array_equal(oq.linspace_device(a, b, n), ref), (a, b, n)


def test_rescale_rules_match_torch_on_the_device(device):
    print("[query] rescale rules vs torch", flush=True)
    rng = np.random.default_rng(1)
    mu = (rng.random(100_000) * 3).astype(np.float32)
    t = torch.from_numpy(mu).to(device)
    ref = ((t / L_MAX) * SPAN + LO).cpu().numpy()
    assert np.array_equal(oq.rescale(mu, L_MAX, SPAN, LO, "reciprocal"), ref)
    h = t.half()
    ref16 = (((h / L_MAX) * SPAN) + LO).float().cpu().numpy()
    assert np.array_equal(oq.rescale(mu, L_MAX, SPAN, LO, "reciprocal_f16"), ref16)
    ref_half_out = ((h.float() / L_MAX) * SPAN + LO).cpu().numpy()
    assert np.array_equal(oq.rescale(mu, L_MAX, SPAN, LO, "reciprocal", mu_dtype="float16"), ref_half_out)
    cpu = ((torch.from_numpy(mu) / L_MAX) * SPAN + LO).numpy()
    assert np.array_equal(oq.rescale(mu, L_MAX, SPAN, LO, "divide"), cpu)


@pytest.mark.parametrize("dtype,dims", [(torch.float32, 3), (torch.float16, 3), (torch.bfloat16, 4),
                                        (torch.float16, 4)])
def test_query_kernel_matches_oracle(device, dtype, dims):
    """Coordinates (in-kernel torch device rule, or explicit axes) and rescale (three modes, network output
    rounded to half or not) bit-exact against the oracle applied to the module path's values; those values
    against the float64 model within the dtype's bar."""
    from nect_amd.serving import QueryRequest, QueryServer

    print(f"[query] kernel vs oracle {dtype} D={dims}", flush=True)
    grid, mlp, w, tab, spec, cfg = model(device, dtype, dims=dims, hidden=3 if dims == 4 else 2)
    # brick traversal (7 x 33 x 40 pads to 8 x 36 x 40 in 4^3 bricks) and row order
    servers = [QueryServer(grid, mlp, w, tab, out_dtype=torch.float32, chunk=3000, max_jobs=5, bricks=b)
               for b in (True, False)]
    n, a, b, t = (7, 33, 40), (0.1, 0.2, 0.05), (0.8, 0.9, 0.95), 0.37
    axes = [oq.linspace_device(a[k], b[k], n[k]) for k in range(3)]
    pts = oq.grid_points(axes, t if dims == 4 else None)
    mu = module_mu(grid, mlp, w, tab, pts, device)
    # the module path's values against the float64 model of the same (rounded) parameters
    params64 = torch.cat([w.double().cpu(), tab.double().cpu()])
    ref64 = model_forward(spec, cfg, params64, pts, round_dtype=None if dtype == torch.float32 else dtype)
    ref64 = ref64.detach().numpy()
    tol = {torch.float32: 2e-5, torch.float16: 6e-3, torch.bfloat16: 4e-2}[dtype]
    assert np.abs(mu - ref64).max() / np.abs(ref64).max() < tol
    for srv in servers:
        for explicit in (False, True):
            for mode in MODES:
                for half in (False, True):
                    kw = dict(l_max=L_MAX, span=SPAN, lo=LO, rescale=oq.RESCALE[mode], mu_half=half)
                    if explicit:
                        kw["axes"] = tuple(torch.from_numpy(ax).to(device) for ax in axes)
                    r = QueryRequest(n, a, b, t, **kw)
                    got = srv.run_alone(r).flatten().cpu().numpy()
                    ref = oq.rescale(mu, L_MAX, SPAN, LO, mode, "float16" if half else "float32")
                    assert np.array_equal(got, ref), (srv.bricks, explicit, mode, half, np.abs(got - ref).max())


def test_mixed_results_equal_alone_at_full_size(device):
    """512^2 slices, a 256^3 sub-volume and a 512^3 volume served together (FP16 out, BF16 4D model): every
    result equals the request served alone, bit for bit."""
    from nect_amd.serving import QueryServer, slice_request, subvolume_request, volume_request

    print("[query] alone vs mixed, 512^2 / 256^3 / 512^3", flush=True)
    grid, mlp, w, tab, _, _ = model(device, torch.bfloat16, dims=4, hidden=3, log2T=21, seed=3)
    reqs = [slice_request("xy", 0.41, 0.2), volume_request(0.6), slice_request("xz", 0.5, 0.9),
            subvolume_request((100, 37, 200), 256, 0.33), slice_request("yz", 0.07, 0.5)]
    for i, r in enumerate(reqs):
        r.arrival = 0.002 * i
        r.rid = i
    srv = QueryServer(grid, mlp, w, tab, chunk=1 << 18, max_jobs=16)
    rows = QueryServer(grid, mlp, w, tab, chunk=1 << 18, max_jobs=16, bricks=False)
    stats = srv.serve(reqs, keep_outputs=True)
    assert stats["requests"] == len(reqs)
    for r in reqs:
        alone = type(r)(r.n, r.a, r.b, r.t, kind=r.kind)
        ref = srv.run_alone(alone).flatten()
        assert torch.equal(r.out, ref), r.kind
        # the brick traversal writes every voxel exactly as row order does
        assert r.brick != 0
        assert torch.equal(rows.run_alone(type(r)(r.n, r.a, r.b, r.t, kind=r.kind)).flatten(), ref), r.kind
        r.out = None


def _static_run(tmp_path, device, mode="static", n_angles=10):
    """A run directory as the trainer writes it: config.yaml (with the projections' range) and
    checkpoints/last.ckpt holding random parameters."""
    from nect_amd.config import get_static_cfg, setup_cfg
    from nect_amd.fused import init_params
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc
    from nect_amd.sampling.geometry import Geometry
    from nect_amd.trainer import checkpoint_state

    geo = Geometry(nDetector=(24, 24), dDetector=(1.0, 1.0), mode="cone", DSD=400.0, DSO=200.0,
                   nVoxel=(24, 20, 16), dVoxel=(0.5, 0.5, 0.5),
                   angles=[2 * math.pi * i / n_angles for i in range(n_angles)])
    cfg = get_static_cfg("hash_grid")
    cfg.update({"mode": mode, "geometry": geo,
                "encoder": {"otype": "HashGrid", "n_levels": 16, "n_features_per_level": 2, "log2_hashmap_size": 14,
                            "base_resolution": 8, "per_level_scale": 1.3},
                "net": {"otype": "FullyFusedMLP", "activation": "LeakyReLU", "output_activation": "Softplus",
                        "n_neurons": 64, "n_hidden_layers": 2},
                "data_range": {"maximum": 2.75, "minimum": 0.31}})
    config = setup_cfg(cfg)
    dims = 4 if mode == "dynamic" else 3
    grid = GridDesc(dims, 16, 2, 14, 8, config.per_level_scale())
    mlp = MlpDesc(32, 1, 64, 2, "LeakyReLU", "Softplus")
    p = init_params(grid, mlp, seed=11, device=device)
    g = torch.Generator().manual_seed(12)
    p[mlp.n_params:] = ((torch.rand(grid.n_params, generator=g) * 2 - 1) * 0.4).to(device)
    base = tmp_path / mode
    (base / "checkpoints").mkdir(parents=True)
    torch.save(checkpoint_state(p, None, None, 0, 0.0), base / "checkpoints" / "last.ckpt")
    config.to_yaml(base / "config.yaml")
    return base, config, grid, mlp, p


def test_export_volume_matches_oracle(device, tmp_path):
    from nect_amd.export import export_volume, read_tiff

    print("[query] export_volume vs oracle", flush=True)
    base, config, grid, mlp, p = _static_run(tmp_path, device)
    sh = p.half()
    for kw in ({}, {"binning": 2, "ROIx": [2, 14], "ROIy": [4, 18]}, {"ROIz": [3, 21], "ROIx": [0, 16]}):
        out = export_volume(base, **kw)
        vol = read_tiff(out / "Static_Volume.tiff")
        axes = oq.static_export_axes(config.geometry.nVoxel, kw.get("binning", 1), kw.get("ROIx"), kw.get("ROIy"),
                                     kw.get("ROIz"))
        assert vol.shape == tuple(a.size for a in axes) and vol.dtype == np.float32
        mu = module_mu(grid, mlp, sh[: mlp.n_params], sh[mlp.n_params:], oq.grid_points(axes), device)
        ref = oq.rescale(mu, config.geometry.max_distance_traveled, 2.75 - 0.31, 0.31, "divide", "float16")
        assert np.array_equal(vol.ravel(), ref), kw


def test_dynamic_exports_match_oracle(device, tmp_path):
    from nect_amd.export import export_volumes, export_volumes_at_timesteps, read_tiff

    print("[query] dynamic exports vs oracle", flush=True)
    base, config, grid, mlp, p = _static_run(tmp_path, device, mode="dynamic", n_angles=10)
    sh = p.half()
    w, tab = sh[: mlp.n_params], sh[mlp.n_params:]
    L = config.geometry.max_distance_traveled
    out = export_volumes_at_timesteps(base, [0, 4, 7], ROIz=[2, 10], ROIy=[3, 17])
    axes = oq.dynamic_export_axes(config.geometry.nVoxel, 1, None, [3, 17], [2, 10])
    for t in oq.export_timesteps([0, 4, 7], 10):
        vol = read_tiff(out / f"T{t:04}.tiff")
        mu = module_mu(grid, mlp, w, tab, oq.grid_points(axes, t), device)
        assert np.array_equal(vol.ravel(), oq.rescale(mu, L, 2.75 - 0.31, 0.31, "reciprocal", "float16")), t
    # per revolution (10 views of one turn), two timesteps averaged: the reference's torch operations
    export_volumes(base, binning=2, avg_timesteps=2)
    axes = oq.dynamic_export_axes(config.geometry.nVoxel, 2)
    lin = torch.linspace(0, 1, 10, device=device)
    for j in (0, 6):
        acc = torch.zeros(tuple(a.size for a in axes), device=device)
        for avg in range(2):
            tt = float((lin[j] + avg / 10).float())
            mu = module_mu(grid, mlp, w, tab, oq.grid_points(axes, tt), device)
            acc += torch.from_numpy(mu).to(device).half().view(acc.shape) / 2
        ref = ((acc / L) * (2.75 - 0.31) + 0.31).cpu().numpy()
        vol = read_tiff(base / "volumesfloat32" / f"T0000_{j:04}.tiff")
        assert np.array_equal(vol, ref), j


def test_create_volume_matches_oracle(device):
    """Trainer.create_volume: voxel centres (device linspace), values rounded to half after each rescale
    operation for a half-precision model (base_trainer.py:402-450)."""
    from nect_amd.config import get_static_cfg, setup_cfg
    from nect_amd.sampling.geometry import Geometry
    from nect_amd.trainer import Trainer

    print("[query] create_volume vs oracle", flush=True)
    n_views = 6
    geo = Geometry(nDetector=(24, 24), dDetector=(1.0, 1.0), mode="cone", DSD=400.0, DSO=200.0,
                   nVoxel=(12, 10, 8), dVoxel=(1.0, 1.0, 1.0),
                   angles=[2 * math.pi * i / n_views for i in range(n_views)])
    cfg = get_static_cfg("hash_grid")
    cfg.update({"epochs": 1, "geometry": geo, "sample_outside": 2,
                "encoder": {"otype": "HashGrid", "n_levels": 16, "n_features_per_level": 2, "log2_hashmap_size": 14,
                            "base_resolution": 8, "per_level_scale": 1.3},
                "net": {"otype": "FullyFusedMLP", "activation": "LeakyReLU", "output_activation": "Softplus",
                        "n_neurons": 64, "n_hidden_layers": 2}})
    proj = torch.rand((n_views, 24, 24), generator=torch.Generator().manual_seed(0)) * 2 + 0.5
    tr = Trainer(setup_cfg(cfg), proj, save_ckpt=False, save_last=False, verbose=False)
    g = torch.Generator().manual_seed(1)
    tr.fused.params[tr.fused.n_mlp:] = ((torch.rand(tr.grid_desc.n_params, generator=g) * 2 - 1) * 0.4).to(device)
    tr.fused.refresh_shadow()
    vol = tr.create_volume(cpu=True).numpy()
    axes = oq.create_volume_axes((12, 10, 8), sample_outside=2)
    mu = module_mu(tr.grid_desc, tr.mlp_desc, tr.fused.mlp_weights, tr.fused.table, oq.grid_points(axes), device)
    span = float(proj.max() - proj.min())
    ref = oq.rescale(mu, geo.max_distance_traveled, span, float(proj.min()), "reciprocal_f16", "float16")
    assert vol.shape == (12, 10, 8)
    assert np.array_equal(vol.ravel(), ref)
