"""GPU parity of the fused training step (sampler -> encoder -> MLP -> integral/loss -> backward) and Adam."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import hashgrid as hg
from oracle.train import train_step
from tests.helpers import cfg1_geometry, product_struct, seeded_batch

pytestmark = pytest.mark.gpu

# line-integral tolerance (relative to the largest projection): the FP32 and FP16/BF16 bars
YTOL = {torch.float32: 1e-5, torch.float16: 2e-3, torch.bfloat16: 2e-3}
GTOL = {torch.float32: 2e-4, torch.float16: 2e-2, torch.bfloat16: 6e-2}


def build(dtype, seed=0, log2T=17, hidden=2, n_levels=16, n_features=2, n_dims=3):
    from nect_amd.fused import FusedTrainStep, init_params
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc

    grid = GridDesc(n_dims, n_levels, n_features, log2T, 16, 1.5)
    mlp = MlpDesc(n_levels * n_features, 1, 64, hidden, "ReLU", "Softplus")
    params = init_params(grid, mlp, seed=seed, device="cuda")
    # a table that is not ~0 so that every term contributes
    g = torch.Generator().manual_seed(seed + 1)
    params[mlp.n_params:] = ((torch.rand(grid.n_params, generator=g) * 2 - 1) * 0.5).cuda()
    fs = FusedTrainStep(grid, mlp, params, dtype, max_rays=2048, max_points_per_ray=64, loss="L2")
    spec = hg.GridSpec(3, 16, 2, log2T, 16, 1.5)
    cfg = {"n_params": mlp.n_params, "n_input_padded": 32, "n_hidden_layers": hidden, "n_output_padded": 16,
           "activation": "ReLU", "output_activation": "Softplus"}
    return fs, spec, cfg


@pytest.mark.parametrize("dtype", [torch.float32, torch.float16, torch.bfloat16])
@pytest.mark.parametrize("loss", ["L2", "L1"])
def test_fused_step_matches_oracle(device, dtype, loss):
    print(f"[train] fused step {dtype} {loss}", flush=True)
    g = cfg1_geometry()
    f = g.c_struct_fields()
    S, B = 64, 768
    fs, spec, cfg = build(dtype)
    from nect_amd.fused import loss_code

    fs.loss = loss_code(loss)
    idx, jit, off = seeded_batch(64 * 64, B, seed=3, det_offset_scale=0.5)
    tgt = np.random.default_rng(4).random(B).astype(np.float32) * 0.3
    step = g.max_distance_traveled / S
    dev = device
    lv = fs.step(product_struct(f), 0.9, torch.from_numpy(idx).to(dev), 0, B, S, step, True,
                 torch.from_numpy(jit).to(dev), torch.from_numpy(off).to(dev), torch.from_numpy(tgt).to(dev),
                 g.max_distance_traveled, lr=0.0, optimizer_step=False)
    torch.cuda.synchronize()
    params_seen = fs.shadow.double().cpu().numpy()
    ref = train_step(f, 0.9, idx, 0, B, S, step, True, jit, off, tgt, g.max_distance_traveled, spec, cfg,
                     params_seen, loss=loss, round_dtype=None if dtype == torch.float32 else dtype)
    y = fs.y_pred[:B].double().cpu().numpy()
    scale = np.abs(ref["y_pred"]).max()
    assert scale > 0.05
    ey = np.abs(y - ref["y_pred"]).max() / scale
    assert ey < YTOL[dtype], ey
    assert abs(float(lv.item()) - ref["loss"]) < 10 * YTOL[dtype] * max(ref["loss"], 1e-6) + 1e-9
    grad = fs.grad.double().cpu().numpy()
    n_mlp = cfg["n_params"]
    for part, sl in (("mlp", slice(0, n_mlp)), ("grid", slice(n_mlp, None))):
        r = ref["grad"][sl]
        e = np.abs(grad[sl] - r).max() / np.abs(r).max()
        assert e < GTOL[dtype], (part, e)


def test_sample_counts_and_mask_match(device):
    """The fused path keeps exactly the samples the reference trainer keeps."""
    from nect_amd.sampling import ct_sampling as cs
    from oracle.sampler import oracle_sample

    g = cfg1_geometry()
    f = g.c_struct_fields()
    S, B = 64, 4096
    fs, _, _ = build(torch.float16)
    idx, jit, off = seeded_batch(64 * 64, B, seed=8)
    dev = device
    fs.forward_rays(product_struct(f), 2.5, torch.from_numpy(idx).to(dev), 0, B, S, g.max_distance_traveled / S,
                    True, torch.from_numpy(jit).to(dev), torch.from_numpy(off).to(dev))
    torch.cuda.synchronize()
    pts, dist, cnt = oracle_sample(f, 2.5, idx, B, 0, S, g.max_distance_traveled / S, True, jit, off)
    _, _, c, d = cs.ray_fields(fs.rays[:B])
    assert np.array_equal(c.cpu().numpy(), cnt)
    to = fs.tile_offsets[: B + 1].cpu().numpy()
    assert int(fs.n_tiles.item()) == int(((cnt + 31) // 32).sum()) == to[-1]


def test_adam_matches_torch(device):
    from nect_amd import _lib

    n = 100003
    g = torch.Generator().manual_seed(0)
    p0 = torch.randn(n, generator=g)
    p_ref = p0.clone().cuda().requires_grad_(True)
    opt = torch.optim.Adam([p_ref], lr=1e-2, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.01)
    p = p0.clone().cuda()
    grad = torch.zeros(n, device=device)
    m = torch.zeros(n, device=device)
    v = torch.zeros(n, device=device)
    shadow = torch.zeros(n, dtype=torch.float16, device=device)
    for step in range(1, 6):
        gr = torch.randn(n, generator=g).cuda() * (10.0 ** -step)
        gr[::7] = 0  # untouched entries still move (dense Adam)
        p_ref.grad = gr.clone()
        opt.step()
        grad.copy_(gr)
        _lib.call("nect_adam", p.data_ptr(), grad.data_ptr(), m.data_ptr(), v.data_ptr(), shadow.data_ptr(), 1, n,
                  1e-2, 0.9, 0.999, 1e-8, 0.01, step, 0.0, _lib.stream_of(p.device))
    torch.cuda.synchronize()
    assert torch.allclose(p, p_ref.detach(), rtol=1e-5, atol=1e-6)
    assert torch.count_nonzero(grad) == 0
    assert torch.equal(shadow, p.half())


def test_fp32_gradient_matches_central_differences(device):
    """Parallel-beam slab, FP32 fused step: the gradient against float64 central differences (h = 1e-3) of
    the loss on every MLP weight and 256 random touched hash entries (line integrals within rel 1e-5)."""
    from nect_amd.fused import FusedTrainStep
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc
    from tests.helpers import FdCase

    print("[train] fp32 gradient vs central differences", flush=True)
    c = FdCase(seed=0)
    grid = GridDesc(3, 16, 2, 20, 16, 1.5)
    mlp = MlpDesc(32, 1, 64, 2, "ReLU", "Softplus")
    assert grid.n_params + mlp.n_params == c.params.size
    params = torch.from_numpy(c.params.astype(np.float32)).to(device)
    fs = FusedTrainStep(grid, mlp, params, torch.float32, max_rays=c.B, max_points_per_ray=c.S, loss="L2")
    lv = fs.step(product_struct(c.fields), c.angle, torch.from_numpy(c.idx).to(device), 0, c.B, c.S, c.step_mm,
                 True, torch.from_numpy(c.jit).to(device), torch.from_numpy(c.off).to(device),
                 torch.from_numpy(c.target).to(device), c.l_max, lr=0.0, optimizer_step=False)
    torch.cuda.synchronize()
    assert abs(float(lv.item()) - c.loss()) < 1e-4 * c.loss()
    which = c.fd_indices()
    fd = c.central_differences(which)
    g = fs.grad.double().cpu().numpy()[which]
    n_mlp = c.cfg["n_params"]
    e_mlp = np.linalg.norm(fd[:n_mlp] - g[:n_mlp]) / np.linalg.norm(fd[:n_mlp])
    e_grid = np.linalg.norm(fd[n_mlp:] - g[n_mlp:]) / np.linalg.norm(fd[n_mlp:])
    print(f"[train] FD rel errors: mlp {e_mlp:.2e}, grid {e_grid:.2e}", flush=True)
    assert e_mlp < 3e-3, e_mlp
    assert e_grid < 2e-2, e_grid


@pytest.mark.parametrize("dtype,grouped", [(torch.float32, False), (torch.float16, False), (torch.float16, True)])
def test_memory_budget_chunks_match_one_pass(device, dtype, grouped):
    """A batch processed in ray chunks under a memory budget gives the loss, line integrals and gradient of
    the one-pass step (up to float summation order); grouped: both on the grouped schedule, each chunk taken
    in patch order."""
    from nect_amd.fused import FusedTrainStep

    print(f"[train] memory-budget chunks {dtype} grouped={grouped}", flush=True)
    g = cfg1_geometry()
    f = g.c_struct_fields()
    S, B = 64, 2048
    fs, _, _ = build(dtype)
    params = fs.params.clone()
    per_ray = fs.slot_bytes() * 2 * 32
    chunked = FusedTrainStep(fs.grid, fs.mlp, params.clone(), dtype, max_rays=B, max_points_per_ray=S, loss="L2",
                             memory_budget=per_ray * 700)
    assert chunked.chunk_rays == 700
    idx, jit, off = seeded_batch(64 * 64, B, seed=6, det_offset_scale=0.5)
    tgt = torch.from_numpy(np.random.default_rng(7).random(B).astype(np.float32) * 0.3).to(device)
    args = (product_struct(f), 1.3, torch.from_numpy(idx).to(device), 0, B, S, g.max_distance_traveled / S, True,
            torch.from_numpy(jit).to(device), torch.from_numpy(off).to(device), tgt, g.max_distance_traveled)
    one = FusedTrainStep(fs.grid, fs.mlp, params.clone(), dtype, max_rays=B, max_points_per_ray=S, loss="L2")
    if grouped:
        one.GROUPS_MIN_WORKGROUPS = chunked.GROUPS_MIN_WORKGROUPS = 0
    l1 = one.step(*args, lr=0.0, optimizer_step=False)
    l2 = chunked.step(*args, lr=0.0, optimizer_step=False)
    torch.cuda.synchronize()
    assert one.last_samples() == chunked.last_samples()
    assert abs(float(l1) - float(l2)) <= 1e-5 * abs(float(l1))
    assert torch.allclose(one.y_pred[:B], chunked.y_pred[:B], rtol=1e-6, atol=0)
    scale = one.grad.abs().max()
    assert (one.grad - chunked.grad).abs().max() <= 1e-4 * scale


@pytest.mark.parametrize("dtype,loss,grouped", [(torch.float32, "L2", False), (torch.float16, "L2", False),
                                                (torch.float32, "L1", False), (torch.float32, "L2", True)])
def test_continuous_scanning_step_matches_oracle(device, dtype, loss, grouped):
    """The continuous-scanning step (mean line integral over K = 3 sub-angles) against the oracle; grouped:
    on the grouped schedule, the batch taken in patch order (per-ray outputs back in the caller's order)."""
    from nect_amd.fused import loss_code
    from oracle.train import train_step_accumulated

    print(f"[train] continuous scanning {dtype} {loss} grouped={grouped}", flush=True)
    g = cfg1_geometry()
    f = g.c_struct_fields()
    S, B, K = 64, 512, 3
    fs, spec, cfg = build(dtype)
    fs.loss = loss_code(loss)
    if grouped:
        fs.GROUPS_MIN_WORKGROUPS = 0
    rng = np.random.default_rng(11)
    idx = rng.permutation(64 * 64).astype(np.int64)
    jits = [(rng.random(B, dtype=np.float32) - np.float32(0.5)) for _ in range(K)]
    offs = [((rng.random((B, 2), dtype=np.float32) - np.float32(0.5)) * np.float32(0.5)) for _ in range(K)]
    lo, hi = 0.6, 0.6 + 2 * np.pi / 90
    edges = np.linspace(lo, hi, K + 1)
    angles = [float((edges[k] + edges[k + 1]) / 2) for k in range(K)]
    tgt = rng.random(B).astype(np.float32) * 0.3
    step = g.max_distance_traveled / S
    lv = fs.step_accumulated(product_struct(f), angles, torch.from_numpy(idx).to(device), 0, B, S, step, True,
                             [torch.from_numpy(j).to(device) for j in jits],
                             [torch.from_numpy(o).to(device) for o in offs], torch.from_numpy(tgt).to(device),
                             g.max_distance_traveled, lr=0.0, optimizer_step=False)
    torch.cuda.synchronize()
    ref = train_step_accumulated(f, angles, idx, 0, B, S, step, True, jits, offs, tgt, g.max_distance_traveled,
                                 spec, cfg, fs.shadow.double().cpu().numpy(), loss=loss,
                                 round_dtype=None if dtype == torch.float32 else dtype)
    y = fs.y_pred[:B].double().cpu().numpy()
    scale = np.abs(ref["y_pred"]).max()
    assert np.abs(y - ref["y_pred"]).max() / scale < YTOL[dtype]
    assert abs(float(lv.item()) - ref["loss"]) < 10 * YTOL[dtype] * max(ref["loss"], 1e-6) + 1e-9
    grad = fs.grad.double().cpu().numpy()
    e = np.abs(grad - ref["grad"]).max() / np.abs(ref["grad"]).max()
    assert e < GTOL[dtype], e


@pytest.mark.parametrize("B,seed,L,F,D", [(4096, 8, 16, 2, 3), (768, 9, 16, 2, 3), (4096, 10, 19, 4, 3),
                                         (1000, 11, 8, 4, 3), (4096, 12, 16, 2, 4)])
def test_grouped_scatter_matches_tile_scatter(device, B, seed, L, F, D):
    """nect_grid_backward_ray_groups (ray groups by detector patch, the coarsest levels summed per 64-byte block
    in LDS, the rest scattered directly) adds the same table gradient as nect_grid_backward_rays, whatever the
    plan and with or without the patch order; the order is a permutation grouping 8x8 patches."""
    from nect_amd import _lib

    print(f"[train] grouped scatter B={B} L={L} F={F} D={D}", flush=True)
    g = cfg1_geometry()
    f = g.c_struct_fields()
    S = 64
    fs, _, _ = build(torch.float32 if L * F <= 64 else torch.float16, n_levels=L, n_features=F, n_dims=D)
    t = 0.37 if D == 4 else 0.0
    fs.scatter = "tiles"
    fs._alloc(B, S)
    idx, jit, off = seeded_batch(64 * 64, B, seed=seed, det_offset_scale=0.5)
    tgt = torch.from_numpy(np.random.default_rng(seed).random(B).astype(np.float32) * 0.3).to(device)
    ridx = torch.from_numpy(idx).to(device)
    geo = product_struct(f)
    fs.step(geo, 0.4, ridx, 0, B, S, g.max_distance_traveled / S, True, torch.from_numpy(jit).to(device),
            torch.from_numpy(off).to(device), tgt, g.max_distance_traveled, lr=0.0, t=t, optimizer_step=False)
    torch.cuda.synchronize()
    n = fs.n_mlp
    ref = fs.grad[n:].clone()
    scale = ref.abs().max()
    assert scale > 0
    fs.scatter = "groups"
    fs.GROUPS_MIN_WORKGROUPS = 0  # small batches take the tile schedule by default
    fs._ray_order(geo, ridx, 0, B, S)
    order = fs.order[:B].cpu().numpy()
    assert np.array_equal(np.sort(order), np.arange(B))
    pix = idx[order]
    row, col = pix // 64, pix % 64
    patch = (row // 8) * 8 + col // 8
    assert np.all(np.diff(patch) >= 0)
    # inside a patch: the 4x4 quarters in turn, each row by row (every pixel is in the batch at most once)
    inner = (((row % 8) // 4) * 2 + (col % 8) // 4) * 16 + (row % 4) * 4 + col % 4
    assert np.all(np.diff(patch * 64 + inner) > 0)
    st = _lib.stream_of(fs.device)
    from nect_amd.fused import aggregation_plan

    auto = aggregation_plan(fs.grid, geo, g.max_distance_traveled / S, B / (64 * 64))
    print(f"[train] plan {auto}", flush=True)
    L = fs.grid.n_levels
    # per-level rounds and one or two launches: none, some, the automatic plan, all levels (the fine ones
    # overflow the LDS table and take the direct fallback), 1 to 16 rounds
    plans = [(list(p) + [0] * L)[:L] for p in ([0], [1] * 3, auto, [1] * 6 + [2, 4, 8, 16], [1] * L, [16] * L,
                                                [-1] * L, [1, 1, 2] + [-1] * 4)]
    for plan in plans:
        for split in (0, 1):
            for use_order in (True, False):
                out = torch.zeros_like(ref)
                _lib.call("nect_grid_backward_ray_groups", fs.grid.struct_ref(), _lib.ptr(fs.rays[:B]),
                          _lib.ptr(fs.tile_offsets), B, _lib.ptr(fs.order) if use_order else None, S, t,
                          _lib.ptr(fs.dfeat), fs.grid.n_output_dims, out.data_ptr(), (ctypes.c_int32 * L)(*plan),
                          split, st)
                torch.cuda.synchronize()
                err = float((out - ref).abs().max() / scale)
                assert err < 1e-5, (plan, split, use_order, err)
    # the fused step on the grouped schedule (its own order and plan; the batch taken in patch order) gives the
    # same table gradient, and the line integrals back in the caller's order
    y_ref, loss_ref = fs.y_pred[:B].clone(), float(fs.loss_value.item())
    for presort in (True, False):
        fs.presort = presort
        fs.grad.zero_()
        fs.y_pred.zero_()
        lv = fs.step(geo, 0.4, ridx, 0, B, S, g.max_distance_traveled / S, True, torch.from_numpy(jit).to(device),
                     torch.from_numpy(off).to(device), tgt, g.max_distance_traveled, lr=0.0, t=t,
                     optimizer_step=False)
        torch.cuda.synchronize()
        assert float((fs.grad[n:] - ref).abs().max() / scale) < 1e-5, presort
        assert torch.allclose(fs.y_pred[:B], y_ref, rtol=1e-6, atol=1e-7), presort
        assert abs(float(lv.item()) - loss_ref) <= 1e-5 * abs(loss_ref), presort


def test_ray_order_is_a_permutation_with_duplicates_and_missing_pixels(device):
    """nect_ray_order on a batch with repeated pixels, pixels outside the detector and slots past the end of
    the index: a permutation; patches in order (rays without a pixel last); inside a patch whose pixels are
    distinct, the 4x4 quarters row by row."""
    from nect_amd import _lib

    print("[train] ray order edge cases", flush=True)
    g = cfg1_geometry()
    geo = product_struct(g.c_struct_fields())
    rng = np.random.default_rng(12)
    n_pix = 64 * 64
    idx = rng.integers(-50, n_pix + 50, size=3000).astype(np.int64)  # repeats, negatives, past the detector
    start, num = 100, 3000  # the last 100 slots read past the index: no pixel
    ridx = torch.from_numpy(idx).to(device)
    need = int(_lib.lib().nect_ray_order_workspace(geo._struct()))
    ws = torch.empty(need, dtype=torch.uint8, device=device)
    order = torch.full((num,), -1, dtype=torch.int32, device=device)
    _lib.call("nect_ray_order", geo._struct(), _lib.ptr(ridx), ridx.numel(), num, start, _lib.ptr(order),
              _lib.ptr(ws), ws.numel(), _lib.stream_of(device))
    torch.cuda.synchronize()
    o = order.cpu().numpy()
    assert np.array_equal(np.sort(o), np.arange(num))
    pix = np.array([idx[start + t] if start + t < idx.size else -1 for t in o])
    valid = (pix >= 0) & (pix < n_pix)
    row, col = np.where(valid, pix // 64, 0), np.where(valid, pix % 64, 0)
    patch = np.where(valid, (row // 8) * 8 + col // 8, 64)
    assert np.all(np.diff(patch) >= 0)
    inner = (((row % 8) // 4) * 2 + (col % 8) // 4) * 16 + (row % 4) * 4 + col % 4
    for p in range(64):
        sel = patch == p
        if sel.sum() > 1 and len(np.unique(pix[sel])) == sel.sum():
            assert np.all(np.diff(inner[sel]) > 0), p


# The two benchmarked training steps at their own shapes: geometry, table size, MLP depth, dtype and the
# scatter's aggregation plan of the full benchmark batch, on a window of neighbouring detector pixels (so that
# the 64-ray groups are whole 8x8 patches and the LDS aggregation really merges samples), batch in random order.
BENCH_STEPS = {
    # 256^3 cone, detector 256^2 @ 1 mm, SOD 400 / SDD 800, hash grid 16x2 T=2^19, MLP 2x64, fp16,
    # <= 256 samples per ray, the plan of a whole-view batch (2^16 rays)
    "cone256_fp16": dict(det=(256, 256), pitch=1.0, D=3, log2T=19, hidden=2, dtype=torch.float16, S=256,
                         window=(32, 32), bench_rays=1 << 16, t=None),
    # 4D model on the 256^3 cone, detector 512^2 @ 0.5 mm, hash grid over (z, y, x, t) 16x2 T=2^21, MLP 3x64,
    # bf16, <= 128 samples per ray, the plan of a 2^17-ray batch, one time t
    "cone4d_bf16": dict(det=(512, 512), pitch=0.5, D=4, log2T=21, hidden=3, dtype=torch.bfloat16, S=128,
                        window=(16, 32), bench_rays=1 << 17, t=0.37),
    # parallel-beam slab 512x512x64 of 1 mm voxels, detector 64 rows x 512 @ 1 mm, hash grid 16x2 T=2^20,
    # MLP 2x64, bf16 (the BF16 side of the slab throughput figures), <= 192 samples per ray, 2^18-ray plan
    "slab_bf16": dict(det=(64, 512), pitch=1.0, D=3, log2T=20, hidden=2, dtype=torch.bfloat16, S=192,
                      window=(16, 32), bench_rays=1 << 18, t=None, mode="parallel", vox=((64, 512, 512), 1.0)),
    # the reference's default static model (cfg/static/hash_grid.yaml) on the 256^3 cone: hash grid 19x4
    # T=2^21, base 8, per-level scale (2*256/8)^(1/18); MLP 6x64 LeakyReLU, ReLU output; L1; fp16; <= 384
    # samples per ray (the end of its sample schedule) and the plan of its 6,510-ray step.  The freshly
    # initialised six-layer network with a ReLU output integrates to at most ~0.012 here, so the
    # non-trivial-output floor is 0.005.
    "static_preset_fp16": dict(det=(256, 256), pitch=1.0, D=3, log2T=21, hidden=6, dtype=torch.float16, S=384,
                               window=(16, 32), bench_rays=6510, t=None, levels=19, features=4, base=8,
                               scale=(2 * 256 / 8) ** (1 / 18), act="LeakyReLU", out_act="ReLU", loss="L1",
                               min_scale=0.005),
}


@pytest.mark.parametrize("name", list(BENCH_STEPS))
def test_benchmarked_step_matches_oracle_at_its_shape(device, name):
    from nect_amd.fused import FusedTrainStep, aggregation_plan, init_params
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc
    from oracle.geometry import OracleGeometry

    c = BENCH_STEPS[name]
    dtype = c["dtype"]
    print(f"[train] benchmarked step {name}", flush=True)
    n_vox, d_vox = c.get("vox", ((256, 256, 256), 0.5))
    if c.get("mode", "cone") == "cone":
        geo = OracleGeometry(mode="cone", nDetector=c["det"], dDetector=(c["pitch"], c["pitch"]), DSD=800.0,
                             DSO=400.0, nVoxel=n_vox, dVoxel=(d_vox,) * 3)
    else:
        geo = OracleGeometry(mode="parallel", nDetector=c["det"], dDetector=(c["pitch"], c["pitch"]), nVoxel=n_vox,
                             dVoxel=(d_vox,) * 3)
    f = geo.c_struct_fields()
    cg = product_struct(f)
    S = c["S"]
    l_max = geo.max_distance_traveled
    step = l_max / S
    L, F = c.get("levels", 16), c.get("features", 2)
    base, scale = c.get("base", 16), c.get("scale", 1.5)
    act, out_act, loss = c.get("act", "ReLU"), c.get("out_act", "Softplus"), c.get("loss", "L2")
    ytol, gtol = c.get("ytol", YTOL[dtype]), c.get("gtol", GTOL[dtype])
    grid = GridDesc(c["D"], L, F, c["log2T"], base, scale)
    mlp = MlpDesc(L * F, 1, 64, c["hidden"], act, out_act)
    params = init_params(grid, mlp, seed=5, device=device)
    gen = torch.Generator().manual_seed(6)
    params[mlp.n_params:] = ((torch.rand(grid.n_params, generator=gen) * 2 - 1) * 0.5).to(device)
    rows, cols = c["det"]
    wr, wc = c["window"]
    B = wr * wc
    fs = FusedTrainStep(grid, mlp, params, dtype, max_rays=B, max_points_per_ray=S, loss=loss)
    fs.GROUPS_MIN_WORKGROUPS = 0          # the large-batch schedule: patch order and grouped scatter
    plan = aggregation_plan(grid, cg, step, c["bench_rays"] / (rows * cols))
    fs.agg_rounds = plan
    print(f"[train] {name}: plan {plan}", flush=True)
    assert any(r > 0 for r in plan) and any(r <= 0 for r in plan)
    rng = np.random.default_rng(7)
    r0, c0 = rows // 2 - wr // 2, cols // 2 - wc // 2 + 8
    pix = ((r0 + np.arange(B) // wc) * cols + (c0 + np.arange(B) % wc)).astype(np.int64)
    idx = rng.permutation(pix)
    jit = (rng.random(B, dtype=np.float32) - np.float32(0.5)).astype(np.float32)
    off = np.zeros((B, 2), np.float32)
    tgt = (rng.random(B) * 0.3).astype(np.float32)
    t = c["t"]
    angle = 1.1
    lv = fs.step(cg, angle, torch.from_numpy(idx).to(device), 0, B, S, step, True, torch.from_numpy(jit).to(device),
                 torch.from_numpy(off).to(device), torch.from_numpy(tgt).to(device), l_max, lr=0.0,
                 t=0.0 if t is None else t, optimizer_step=False)
    torch.cuda.synchronize()
    n_kept = fs.last_samples()
    print(f"[train] {name}: {n_kept} kept samples", flush=True)
    assert n_kept > B * S // 4
    spec = hg.GridSpec(c["D"], L, F, c["log2T"], base, scale)
    cfg = {"n_params": mlp.n_params, "n_input_padded": mlp.n_input_padded, "n_hidden_layers": c["hidden"],
           "n_output_padded": mlp.n_output_padded, "activation": act, "output_activation": out_act}
    ref = train_step(f, angle, idx, 0, B, S, step, True, jit, off, tgt, l_max, spec, cfg,
                     fs.shadow.double().cpu().numpy(), loss=loss, t=t, round_dtype=dtype)
    assert int(ref["counts"].sum()) == n_kept
    y = fs.y_pred[:B].double().cpu().numpy()
    scale = np.abs(ref["y_pred"]).max()
    assert scale > c.get("min_scale", 0.05)
    ey = np.abs(y - ref["y_pred"]).max() / scale
    print(f"[train] {name}: y rel {ey:.2e}", flush=True)
    assert ey < ytol, ey
    assert abs(float(lv.item()) - ref["loss"]) < 10 * ytol * max(ref["loss"], 1e-6) + 1e-9
    grad = fs.grad.double().cpu().numpy()
    n_mlp = cfg["n_params"]
    for part, sl in (("mlp", slice(0, n_mlp)), ("grid", slice(n_mlp, None))):
        r = ref["grad"][sl]
        e = np.abs(grad[sl] - r).max() / np.abs(r).max()
        print(f"[train] {name}: {part} grad rel {e:.2e}", flush=True)
        assert e < gtol, (part, e)


def test_fp16_step_skips_non_finite_gradient_and_backs_off(device):
    """precision="16-mixed" semantics on the device: a batch whose gradient is not finite (an Inf target) leaves
    params, m, v and the fp16 copy bit-unchanged, zeroes the gradient, halves the loss scale and does not count
    as an Adam step; finite steps count, and the scale doubles after amp_growth_interval of them in a row."""
    print("[train] fp16 overflow guard", flush=True)
    g = cfg1_geometry()
    f = g.c_struct_fields()
    S, B = 64, 1024
    fs, _, _ = build(torch.float16)
    fs.amp_growth_interval = 3
    idx, jit, off = seeded_batch(64 * 64, B, seed=13, det_offset_scale=0.5)
    tgt = torch.from_numpy(np.random.default_rng(14).random(B).astype(np.float32) * 0.3).to(device)
    step = g.max_distance_traveled / S
    geo = product_struct(f)

    def run(target):
        return fs.step(geo, 0.3, torch.from_numpy(idx).to(device), 0, B, S, step, True,
                       torch.from_numpy(jit).to(device), torch.from_numpy(off).to(device), target,
                       g.max_distance_traveled, lr=1e-3)

    run(tgt)
    st0 = fs.amp_status()
    assert st0["step"] == 1 and st0["skipped"] == 0 and st0["found_inf"] == 0
    assert st0["scale"] == 2.0 ** round(np.log2(0.5 * B * S))
    snap = [t.clone() for t in (fs.params, fs.m, fs.v, fs.shadow)]
    bad = tgt.clone()
    bad[5] = float("inf")
    lv = run(bad)
    torch.cuda.synchronize()
    assert not np.isfinite(float(lv.item()))
    st1 = fs.amp_status()
    assert st1["step"] == 1 and st1["skipped"] == 1 and st1["found_inf"] == 0
    assert st1["scale"] == st0["scale"] / 2
    for a, b in zip((fs.params, fs.m, fs.v, fs.shadow), snap):
        assert torch.equal(a.view(torch.int16) if a.dtype == torch.float16 else a.view(torch.int32),
                           b.view(torch.int16) if b.dtype == torch.float16 else b.view(torch.int32))
    assert torch.count_nonzero(fs.grad) == 0
    # finite steps again: counted; the third in a row doubles the scale
    for k in range(3):
        run(tgt)
    st2 = fs.amp_status()
    assert st2["step"] == 4 and st2["skipped"] == 1
    assert st2["scale"] == st1["scale"] * 2
    assert not torch.equal(fs.params, snap[0])
    assert torch.isfinite(fs.params).all() and torch.isfinite(fs.shadow.float()).all()
