"""GPU tests of the fused query kernel and the continuous-batching server."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def model(device, dtype, dims=4, hidden=3, log2T=19, seed=0):
    from nect_amd.fused import init_params
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc

    grid = GridDesc(dims, 16, 2, log2T, 16, 1.5)
    mlp = MlpDesc(32, 1, 64, hidden, "LeakyReLU", "None")
    p = init_params(grid, mlp, seed=seed, device=device)
    g = torch.Generator().manual_seed(seed + 7)
    p[mlp.n_params:] = ((torch.rand(grid.n_params, generator=g) * 2 - 1) * 0.3).to(device)
    shadow = p.to(dtype)
    return grid, mlp, shadow[: mlp.n_params], shadow[mlp.n_params:]


def reference_values(grid, mlp, weights, table, r, device):
    """Module-path evaluation on torch.linspace/meshgrid coordinates, as the reference export builds them."""
    from nect_amd import grid as G
    from nect_amd import mlp as M

    axes = [torch.linspace(r.a[k], r.b[k], steps=r.n[k], device=device) for k in range(3)]
    z, y, x = torch.meshgrid(axes, indexing="ij")
    cols = [z.flatten(), y.flatten(), x.flatten()]
    if grid.n_dims == 4:
        cols.append(torch.full_like(cols[0], r.t))
    pts = torch.stack(cols, 1).contiguous()
    feats = G.encode(grid, table, pts)
    return M.forward(mlp, weights, feats)[:, 0]


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16, torch.float32])
def test_query_kernel_matches_module_path(device, dtype):
    from nect_amd.serving import QueryServer, slice_request, subvolume_request

    print(f"[serving] query vs module {dtype}", flush=True)
    grid, mlp, w, tab = model(device, dtype)
    srv = QueryServer(grid, mlp, w, tab, out_dtype=torch.float32, chunk=5000, max_jobs=4)
    for r in [slice_request("xy", 0.37, 0.25, size=96), slice_request("yz", 0.9, 0.5, size=64),
              subvolume_request((3, 10, 20), 40, 0.75, full=128)]:
        got = srv.run_alone(r).flatten()
        ref = reference_values(grid, mlp, w, tab, r, device)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (r.kind, (got - ref).abs().max().item())


def test_linspace_coordinates_match_torch(device):
    """In-kernel coordinates equal torch.linspace on the device: a 1-level dense grid with an affine table
    reproduces the coordinate exactly, so the output is the coordinate."""
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc
    from nect_amd.serving import QueryRequest, QueryServer

    # identity network on feature 0: W0 picks feature 0, ReLU hidden, output weight 1
    grid = GridDesc(3, 16, 2, 19, 16, 1.5)
    mlp = MlpDesc(32, 1, 64, 1, "None", "None")
    w = torch.zeros(mlp.n_params, device=device)
    w[0] = 1.0                      # W0[0][0] = 1 (hidden 0 <- feature 0)
    w[64 * 32] = 1.0                # Wout[0][0] = 1
    table = torch.zeros(grid.n_params, device=device)
    # level 0 is dense with resolution 16: feature 0 of entry (c0, c1, c2) = c0 (the z corner)
    res = grid.resolution[0]
    e = torch.arange(grid.level_size[0], device=device)
    table[0: 2 * grid.level_size[0]: 2] = (e % res).float()
    srv = QueryServer(grid, mlp, w, table, out_dtype=torch.float32)
    r = QueryRequest((37, 1, 1), (0.013, 0.5, 0.5), (0.911, 0.5, 0.5), 0.0)
    got = srv.run_alone(r).flatten()
    z = torch.linspace(0.013, 0.911, 37, device=device)
    p = (z * grid.scale[0]) + 0.5   # multilinear interpolation of an affine table reproduces p - 0 exactly
    torch.cuda.synchronize()
    assert torch.allclose(got, p, rtol=0, atol=1e-5), (got - p).abs().max().item()


def test_mixed_results_equal_alone(device):
    from nect_amd.serving import QueryServer, mixed_workload

    print("[serving] alone vs mixed", flush=True)
    grid, mlp, w, tab = model(device, torch.bfloat16)
    reqs = mixed_workload(rate=400.0, duration=0.05, seed=3, size=128)
    assert len(reqs) >= 8
    srv = QueryServer(grid, mlp, w, tab, chunk=1 << 14, max_jobs=8)
    stats = srv.serve(reqs, keep_outputs=True)
    assert stats["requests"] == len(reqs)
    for r in reqs:
        assert r.done is not None and r.latency >= 0
        alone = type(r)(r.n, r.a, r.b, r.t, kind=r.kind)
        ref = srv.run_alone(alone).flatten()
        assert torch.equal(r.out, ref), r.kind


def test_budget_admission_serialises_large_requests(device):
    from nect_amd.serving import QueryServer, volume_request

    grid, mlp, w, tab = model(device, torch.float16, dims=3, hidden=2)
    vols = [volume_request(0.0, size=64) for _ in range(3)]
    for i, v in enumerate(vols):
        v.arrival = 0.0
        v.rid = i
    # budget for one output at a time
    srv = QueryServer(grid, mlp, w, tab, hbm_budget=64 ** 3 * 2, chunk=1 << 15)
    stats = srv.serve(vols)
    assert stats["requests"] == 3
    assert vols[1].admitted >= vols[0].done - 1e-6 and vols[2].admitted >= vols[1].done - 1e-6


def test_host_copy_delivers_results_and_launches_are_timed(device):
    """With host_copy a request completes once its result is in pinned host memory: every delivered host
    buffer equals the device result; with timing every launch has a positive duration and its positions."""
    from nect_amd.serving import QueryServer, slice_request, subvolume_request

    grid, mlp, w, tab = model(device, torch.bfloat16)
    reqs = [slice_request("xy", 0.3, 0.1, size=64), subvolume_request((3, 5, 7), 37, 0.2, full=64),
            slice_request("yz", 0.8, 0.7, size=64)]
    for i, r in enumerate(reqs):
        r.arrival = 0.001 * i
        r.rid = i
    srv = QueryServer(grid, mlp, w, tab, chunk=1 << 12, max_jobs=4, host_copy=True, timing=True)
    stats = srv.serve(reqs, keep_outputs=True)
    assert stats["requests"] == len(reqs)
    for r in reqs:
        assert r.host is not None and r.host.is_pinned()
        assert torch.equal(r.host, r.out.cpu())
    launches = srv.kernel_ms()
    assert len(launches) == srv.launches and all(ms > 0 for ms, _ in launches)
    assert sum(n for _, n in launches) == sum(r.slots for r in reqs)
