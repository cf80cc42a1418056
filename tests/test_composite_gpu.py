"""GPU parity of the composite models (QuadCubes, HyperCubes, a grid with Identity) and wide networks: the
one-launch nested-grid encoder against the per-grid encoder, the modules and their gradients against the
float64 oracle (oracle/networks.py), the module training step against the oracle's step, and the dynamic
reconstruction with the reference's default QuadCubes preset end to end."""
import math

import numpy as np
import pytest
import torch

from oracle import hashgrid as hg
from oracle import networks as on
from oracle.sampler import oracle_sample
from oracle.train import integrate
from tests.helpers import cfg1_geometry, product_struct, seeded_batch

pytestmark = pytest.mark.gpu

ENC = {"otype": "HashGrid", "n_levels": 6, "n_features_per_level": 4, "log2_hashmap_size": 13, "base_resolution": 8,
       "per_level_scale": 1.6}
NET = {"otype": "FullyFusedMLP", "activation": "LeakyReLU", "output_activation": "None", "n_neurons": 128,
       "n_hidden_layers": 3}
YTOL = {torch.float32: 2e-5, torch.float16: 6e-3, torch.bfloat16: 4e-2}
GTOL = {torch.float32: 2e-4, torch.float16: 3e-2, torch.bfloat16: 8e-2}


def _oracle_parts(model_kind: str, identity: bool):
    s3 = hg.GridSpec(3, ENC["n_levels"], 4, ENC["log2_hashmap_size"], ENC["base_resolution"], ENC["per_level_scale"])
    s4 = hg.GridSpec(4, ENC["n_levels"], 4, ENC["log2_hashmap_size"], ENC["base_resolution"], ENC["per_level_scale"])
    parts = [("grid", s3)] * 4
    if model_kind == "hyper":
        parts = parts + [("grid", s4)]
    if identity:
        parts = parts + [("identity", 4)]
    return parts


def _oracle_cfg(net):
    d = net.desc
    return {"n_params": d.n_params, "n_input_padded": d.n_input_padded, "width": d.width,
            "n_hidden_layers": d.n_hidden_layers, "n_output_padded": d.n_output_padded, "activation": d.activation,
            "output_activation": d.output_activation}


def _make(kind, dtype, device, identity=True, seed=0):
    from nect_amd.networks import HyperCubes, QuadCubes

    cls = QuadCubes if kind == "quad" else HyperCubes
    m = cls(ENC, NET, include_identity=identity, dtype=dtype, seed=seed).to(device)
    g = torch.Generator().manual_seed(seed + 1)
    n_net = m.net.desc.n_params
    with torch.no_grad():
        m.net.params[n_net:] = ((torch.rand(m.net.n_enc_params, generator=g) * 2 - 1) * 0.5).to(device)
    return m


def test_parts_encoder_equals_per_grid_encoder(device):
    """nect_grid_encode_parts (four 3-D grids over their column triples, one launch) writes each grid's
    features bit for bit as nect_grid_encode does, and nect_grid_backward_parts adds each grid's gradient."""
    from nect_amd import grid as G
    from nect_amd.tcnn import Encoding

    print("[composite] parts encoder vs per-grid", flush=True)
    grids = [{"n_dims_to_encode": 3, **ENC} for _ in range(4)]
    # rows of 96 columns take the staged write-back; a 3-column identity part makes rows of 99, whose
    # level chunks are not aligned, and the unstaged kernel
    for nested, dtype in [(grids, torch.float32), (grids, torch.float16), (grids, torch.bfloat16),
                          (grids + [{"otype": "Identity", "n_dims_to_encode": 3}], torch.float16),
                          (grids + [{"otype": "Identity", "n_dims_to_encode": 3}], torch.float32)]:
        n_in = 3 * len(nested)
        enc = Encoding(n_in, {"otype": "Composite", "nested": nested}, dtype=dtype).to(device)
        with torch.no_grad():
            enc.params.uniform_(-0.5, 0.5)
        x = torch.rand(3001, n_in, device=device)
        out = enc(x)
        tables = enc.params.to(dtype)
        for part in enc.parts:
            if part.desc is None:  # identity columns: the input as it is
                assert torch.equal(out[:, part.out0:part.out0 + part.n_out].float(),
                                   x[:, part.col0:part.col0 + part.n_dims].to(dtype).float())
                continue
            ref = G.encode(part.desc, tables[part.p0:part.p0 + part.n_params],
                           x[:, part.col0:part.col0 + 3].contiguous())
            assert torch.equal(out[:, part.out0:part.out0 + part.n_out], ref)
        dout = torch.randn_like(out, dtype=torch.float32)
        (out.float() * dout).sum().backward()
        # autograd hands the encoding dL/dout in the output's dtype
        dout = dout.to(dtype).float()
        for part in enc.parts:
            if part.desc is None:
                continue
            ref = torch.zeros(part.n_params, device=device)
            G.backward(part.desc, x[:, part.col0:part.col0 + 3].contiguous(),
                       dout[:, part.out0:part.out0 + part.n_out].contiguous(), ref)
            got = enc.params.grad[part.p0:part.p0 + part.n_params]
            assert float((got - ref).abs().max()) <= 1e-5 * float(ref.abs().max())


@pytest.mark.parametrize("kind,dtype", [("quad", torch.float32), ("quad", torch.float16), ("hyper", torch.bfloat16),
                                        ("hyper", torch.float32)])
def test_composite_module_matches_oracle(device, kind, dtype):
    print(f"[composite] {kind} module vs oracle {dtype}", flush=True)
    m = _make(kind, dtype, device)
    rng = np.random.default_rng(3)
    zyx = rng.random((2000, 3), dtype=np.float32)
    t = 0.37
    y = m(torch.from_numpy(zyx).to(device), t)[:, 0]
    dy = torch.from_numpy(rng.standard_normal(2000).astype(np.float32)).to(device)
    (y * dy).sum().backward()
    torch.cuda.synchronize()
    x = on.quad_inputs(zyx, t) if kind == "quad" else on.hyper_inputs(zyx, t)
    p_seen = m.net.params.detach().to(dtype).double().cpu().requires_grad_(True)
    ref = on.composite_forward(_oracle_parts(kind, True), _oracle_cfg(m.net), p_seen, x,
                               round_dtype=None if dtype == torch.float32 else dtype)
    (ref * dy.double().cpu()).sum().backward()
    got = y.detach().double().cpu()
    ey = float((got - ref.detach()).abs().max() / ref.detach().abs().max())
    print(f"[composite] y rel {ey:.2e}", flush=True)
    assert ey < YTOL[dtype], ey
    g, r = m.net.params.grad.double().cpu(), p_seen.grad
    n_net = m.net.desc.n_params
    for part, sl in (("mlp", slice(0, n_net)), ("grids", slice(n_net, None))):
        e = float((g[sl] - r[sl]).abs().max() / r[sl].abs().max())
        print(f"[composite] {part} grad rel {e:.2e}", flush=True)
        assert e < GTOL[dtype], (part, e)


@pytest.mark.parametrize("dtype,groups", [(torch.float32, False), (torch.float16, False), (torch.float32, True),
                                          (torch.float16, True), (torch.bfloat16, True)])
def test_module_step_matches_oracle(device, dtype, groups):
    """One training step of a QuadCubes model on the module path (rays, tiles, positions, the module with
    autograd, line integral and L2 loss) against the oracle's step: line integrals and gradient; with
    ``groups`` through the patch-ordered batch and the ray-group scatter of each grid."""
    from nect_amd.module_step import ModuleTrainStep

    print(f"[composite] module step vs oracle {dtype} groups={groups}", flush=True)
    m = _make("quad", dtype, device, seed=4)
    g = cfg1_geometry()
    f = g.c_struct_fields()
    S, B, t = 64, 512, 0.61
    step = ModuleTrainStep(m, dtype, B, S, loss="L2", dynamic=True)
    step.scatter = "groups" if groups else "module"
    step.GROUPS_MIN_WORKGROUPS = 0
    idx, jit, off = seeded_batch(64 * 64, B, seed=5, det_offset_scale=0.5)
    tgt = (np.random.default_rng(6).random(B) * 0.3).astype(np.float32)
    lv = step.step(product_struct(f), 0.8, torch.from_numpy(idx).to(device), 0, B, S, g.max_distance_traveled / S,
                   True, torch.from_numpy(jit).to(device), torch.from_numpy(off).to(device),
                   torch.from_numpy(tgt).to(device), g.max_distance_traveled, lr=0.0, t=t, optimizer_step=False)
    torch.cuda.synchronize()
    pts, dist, cnt = oracle_sample(f, 0.8, idx, B, 0, S, g.max_distance_traveled / S, True, jit, off)
    flat = pts.reshape(-1, 3)
    keep = ~np.all(flat == 0, axis=1)
    p_seen = m.net.params.detach().to(dtype).double().cpu().requires_grad_(True)
    mu = on.composite_forward(_oracle_parts("quad", True), _oracle_cfg(m.net), p_seen,
                              on.quad_inputs(flat[keep], t), round_dtype=None if dtype == torch.float32 else dtype)
    full = torch.zeros(B * S, dtype=torch.float64)
    full[torch.from_numpy(keep)] = mu
    y = integrate(full.view(B, S), dist, g.max_distance_traveled)
    L = torch.nn.functional.mse_loss(y, torch.from_numpy(tgt.astype(np.float64)))
    L.backward()
    got_y = step.y_pred[:B].double().cpu()
    ey = float((got_y - y.detach()).abs().max() / y.detach().abs().max())
    assert ey < {torch.float32: 1e-5, torch.float16: 2e-3, torch.bfloat16: 1e-2}[dtype], ey
    assert abs(float(lv.item()) - float(L.detach())) < 1e-2 * float(L.detach())
    gg, r = step.grad.double().cpu(), p_seen.grad
    e = float((gg - r).abs().max() / r.abs().max())
    assert e < GTOL[dtype], e


def test_module_step_compute_copy_is_the_cast_of_the_parameters(device):
    """The FP16 module step hands the module the parameters in FP16 from a copy Adam writes as it updates them
    (no cast of every parameter per step): after every step -- applied, skipped for a non-finite gradient, or
    following a change to the parameters made outside the step -- the copy the next forward uses equals
    params.to(float16) bit for bit."""
    from nect_amd.module_step import ModuleTrainStep

    print("[composite] module step compute copy", flush=True)
    dtype = torch.float16
    m = _make("quad", dtype, device, seed=4)
    g = cfg1_geometry()
    f = g.c_struct_fields()
    S, B, t = 64, 512, 0.61
    st = ModuleTrainStep(m, dtype, B, S, loss="L2", dynamic=True)
    st.GROUPS_MIN_WORKGROUPS = 0
    idx, jit, off = seeded_batch(64 * 64, B, seed=5, det_offset_scale=0.5)
    tgt = torch.from_numpy((np.random.default_rng(6).random(B) * 0.3).astype(np.float32)).to(device)
    bad = tgt.clone()
    bad[3] = float("inf")

    def run(target):
        st.step(product_struct(f), 0.8, torch.from_numpy(idx).to(device), 0, B, S, g.max_distance_traveled / S,
                True, torch.from_numpy(jit).to(device), torch.from_numpy(off).to(device), target,
                g.max_distance_traveled, lr=1e-3, t=t)
        torch.cuda.synchronize()

    def copy_is_cast():
        cp = st._compute_params()
        return cp is not None and torch.equal(cp, m.net.params.detach().to(dtype))

    for target in (tgt, tgt, bad, tgt):
        before = m.net.params.detach().clone()
        run(target)
        assert st._compute_key == st._params_key()
        assert torch.equal(st._compute_copy, m.net.params.detach().to(dtype))
        if target is bad:
            assert torch.equal(m.net.params.detach(), before)      # the skipped step changed nothing
        else:
            assert not torch.equal(m.net.params.detach(), before)
    assert st.amp_status()["skipped"] == 1
    with torch.no_grad():
        m.net.params[:100] += 0.01          # outside the step
    assert st._compute_key != st._params_key()
    assert copy_is_cast()                   # refreshed by a cast before the next forward
    run(tgt)
    assert torch.equal(st._compute_copy, m.net.params.detach().to(dtype))


@pytest.mark.parametrize("kind", ["quad", "hyper"])
def test_module_step_at_the_preset_layers_matches_oracle(device, kind):
    """The dynamic presets' layer structure (cfg/dynamic/quadcubes.yaml, hypercubes.yaml: four 3-D grids -- and
    for HyperCubes a 4-D one -- of 23 levels x 4 features, base 16, per-level scale (2*256/16)^(1/22), identity;
    4 x 128 LeakyReLU, no output activation; 372 or 464 network inputs, the staged encoder rows, the wide first
    layer's products) with the tables cut to T = 2^15 so that the float64 oracle stays small: one FP16 step
    through the grouped schedule against the oracle's step, with the module tests' FP16 tolerances."""
    from nect_amd.module_step import ModuleTrainStep
    from nect_amd.networks import HyperCubes, QuadCubes

    print(f"[composite] {kind} preset layers, module step vs oracle", flush=True)
    scale = (2 * 256 / 16) ** (1 / 22)
    enc = {"otype": "HashGrid", "n_levels": 23, "n_features_per_level": 4, "log2_hashmap_size": 15,
           "base_resolution": 16, "per_level_scale": scale}
    net = {"otype": "FullyFusedMLP", "activation": "LeakyReLU", "output_activation": "None", "n_neurons": 128,
           "n_hidden_layers": 4}
    dtype = torch.float16
    m = (QuadCubes if kind == "quad" else HyperCubes)(enc, net, include_identity=True, dtype=dtype,
                                                       seed=9).to(device)
    gen = torch.Generator().manual_seed(10)
    n_net = m.net.desc.n_params
    with torch.no_grad():
        m.net.params[n_net:] = ((torch.rand(m.net.n_enc_params, generator=gen) * 2 - 1) * 0.5).to(device)
    g = cfg1_geometry()
    f = g.c_struct_fields()
    S, B, t = 64, 512, 0.27
    step = ModuleTrainStep(m, dtype, B, S, loss="L2", dynamic=True)
    step.GROUPS_MIN_WORKGROUPS = 0
    idx, jit, off = seeded_batch(64 * 64, B, seed=11, det_offset_scale=0.5)
    tgt = (np.random.default_rng(12).random(B) * 0.3).astype(np.float32)
    lv = step.step(product_struct(f), 0.4, torch.from_numpy(idx).to(device), 0, B, S, g.max_distance_traveled / S,
                   True, torch.from_numpy(jit).to(device), torch.from_numpy(off).to(device),
                   torch.from_numpy(tgt).to(device), g.max_distance_traveled, lr=0.0, t=t, optimizer_step=False)
    torch.cuda.synchronize()
    pts, dist, cnt = oracle_sample(f, 0.4, idx, B, 0, S, g.max_distance_traveled / S, True, jit, off)
    flat = pts.reshape(-1, 3)
    keep = ~np.all(flat == 0, axis=1)
    parts = [("grid", hg.GridSpec(3, 23, 4, 15, 16, scale))] * 4
    if kind == "hyper":
        parts = parts + [("grid", hg.GridSpec(4, 23, 4, 15, 16, scale))]
    parts = parts + [("identity", 4)]
    inputs = on.quad_inputs(flat[keep], t) if kind == "quad" else on.hyper_inputs(flat[keep], t)
    p_seen = m.net.params.detach().to(dtype).double().cpu().requires_grad_(True)
    mu = on.composite_forward(parts, _oracle_cfg(m.net), p_seen, inputs, round_dtype=dtype)
    full = torch.zeros(B * S, dtype=torch.float64)
    full[torch.from_numpy(keep)] = mu
    y = integrate(full.view(B, S), dist, g.max_distance_traveled)
    L = torch.nn.functional.mse_loss(y, torch.from_numpy(tgt.astype(np.float64)))
    L.backward()
    got_y = step.y_pred[:B].double().cpu()
    ey = float((got_y - y.detach()).abs().max() / y.detach().abs().max())
    print(f"[composite] {kind} preset layers: y rel {ey:.2e}", flush=True)
    assert ey < YTOL[dtype], ey
    gg, r = step.grad.double().cpu(), p_seen.grad
    for part, sl in (("network", slice(0, n_net)), ("grids", slice(n_net, None))):
        e = float((gg[sl] - r[sl]).abs().max() / r[sl].abs().max())
        print(f"[composite] {kind} preset layers: {part} grad rel {e:.2e}", flush=True)
        assert e < GTOL[dtype], (part, e)


@pytest.mark.parametrize("kind,dtype", [("quad", torch.float32), ("hyper", torch.float32), ("quad", torch.float16)])
def test_grouped_composite_scatter_matches_module_scatter(device, kind, dtype):
    """The ray-group scatter of every nested grid (each reading its own selection of z, y, x, t), with the batch
    presorted or not and under several per-level plans (LDS rounds, across rays, direct), gives the gradient
    the module's per-sample scatter gives (float addition order aside), and the same line integrals in the
    caller's order.  FP16: the ray-group scatter reads the network backward's FP16 loss-scaled dL/dfeatures
    and unscales them as it reads; the module's scatter takes the FP32 unscaled copy."""
    from nect_amd.module_step import ModuleTrainStep
    from nect_amd.sampling.geometry import Geometry

    print(f"[composite] grouped vs module scatter {kind} {dtype}", flush=True)
    m = _make(kind, dtype, device, seed=7)
    geo = Geometry(nDetector=(64, 64), dDetector=(1.0, 1.0), mode="cone", DSD=1000.0, DSO=500.0,
                   nVoxel=(64, 64, 64), dVoxel=(0.5, 0.5, 0.5), angles=[0.3])
    cg = geo.get_c_geometry()
    S, B, t = 64, 2048, 0.43
    gen = torch.Generator(device=device).manual_seed(3)
    idx = torch.randperm(64 * 64, device=device, generator=gen)
    jit = torch.rand(B, device=device, generator=gen) - 0.5
    off = (torch.rand((B, 2), device=device, generator=gen) - 0.5) * 0.4
    tgt = torch.rand(B, device=device, generator=gen) * 0.3
    l_max = geo.max_distance_traveled

    def run(scatter, presort=True, rounds=None, split=0):
        st = ModuleTrainStep(m, dtype, B, S, loss="L2", dynamic=True)
        st.scatter, st.presort, st.GROUPS_MIN_WORKGROUPS = scatter, presort, 0
        st.scatter_split = split
        if rounds is not None:
            st.agg_rounds = [(rounds if isinstance(rounds, list) else [rounds] * p.desc.n_levels)
                             for p in m.net.grid_parts()]
        st.step(cg, 0.3, idx, 0, B, S, l_max / S, True, jit, off, tgt, l_max, lr=0.0, t=t, optimizer_step=False)
        torch.cuda.synchronize()
        return st.y_pred[:B].clone(), st.grad.clone()

    y0, g0 = run("module")
    n_net = m.net.desc.n_params
    # split 1: the aggregated and the direct levels as two launches
    mixed = [1, 1, 2, -1, -1, -1]
    for presort, rounds, split in [(True, None, 0), (False, None, 0), (True, 1, 0), (True, 2, 0), (True, -1, 0),
                                   (False, 0, 0), (True, 16, 0), (True, mixed, 0), (True, mixed, 1),
                                   (False, [4, 0, 0, -1, -1, -1], 1), (True, None, 1)]:
        y, gr = run("groups", presort, rounds, split)
        assert torch.equal(y, y0), (presort, rounds, split)
        for sl in (slice(0, n_net), slice(n_net, None)):
            err = float((gr[sl] - g0[sl]).abs().max() / g0[sl].abs().max())
            assert err < 1e-5, (presort, rounds, split, err)


def test_dynamic_reconstruct_trains_the_quadcubes_preset(device, tmp_path, monkeypatch):
    """reconstruct(mode="dynamic") trains the reference's default dynamic model (QuadCubes with its preset
    network, 4 x 128 LeakyReLU, identity columns), here with a 2^14 table, on the projections of a sphere
    that grows with t; the run directory holds the checkpoint and config, and export_volumes_at_timesteps reads
    it back: the exported volume is finite and brighter in the sphere than outside."""
    from nect_amd import phantom
    from nect_amd.export import export_volumes_at_timesteps, read_tiff
    from nect_amd.reconstruct import reconstruct
    from nect_amd.sampling.geometry import Geometry

    print("[composite] dynamic reconstruct, QuadCubes preset", flush=True)
    monkeypatch.chdir(tmp_path)
    n_views = 24
    angles = [2 * math.pi * 2 * i / n_views for i in range(n_views)]
    geo = Geometry(nDetector=(32, 32), dDetector=(1.0, 1.0), mode="cone", DSD=400.0, DSO=200.0,
                   nVoxel=(32, 32, 32), dVoxel=(0.5, 0.5, 0.5), angles=angles)
    projs = []
    for i, a in enumerate(angles):
        r = 3.0 + 2.0 * i / (n_views - 1)
        ph = phantom.Phantom(torch.tensor([[0.0, 0.0, 0.0]], dtype=torch.float64),
                             torch.tensor([[r, r, r]], dtype=torch.float64),
                             torch.eye(3, dtype=torch.float64)[None], torch.tensor([1.0], dtype=torch.float64))
        projs.append(phantom.project(ph, "cone", (32, 32), (1.0, 1.0), 200.0, 400.0, [a], device=device)[0].cpu())
    proj = torch.stack(projs)
    enc = {"otype": "HashGrid", "n_levels": 23, "n_features_per_level": 4, "log2_hashmap_size": 14,
           "base_resolution": 16, "max_resolution_factor": 2}
    run = reconstruct(geo, proj / proj.max(), mode="dynamic", quality="poor", niter=6, verbose=False,
                      config_override={"encoder": enc, "lr": 2e-3,
                                       "warmup": {"steps": 10, "lr0": 0.1, "otype": "exponential"}})
    assert (run / "checkpoints" / "last.ckpt").exists() and (run / "config.yaml").exists()
    out = export_volumes_at_timesteps(run, [0.5], binning=1)
    vol = read_tiff(out / f"T{0.5:04}.tiff")
    assert vol.shape == (32, 32, 32) and np.all(np.isfinite(vol))
    c = np.arange(32) - 15.5
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    rr = np.sqrt(zz ** 2 + yy ** 2 + xx ** 2) * 0.5
    assert vol[rr < 2.0].mean() > vol[rr > 6.0].mean()


@pytest.mark.parametrize("mode", ["reciprocal", "divide"])
def test_module_export_matches_oracle(device, mode):
    """Exports of a composite model (ModuleExportModel: the module in chunks, output in the compute dtype,
    the reference's torch rescale): values against the float64 oracle within the fp16 bar, and the rescale
    bit for bit against the oracle's rule applied to the module's own output."""
    from types import SimpleNamespace

    from nect_amd.export import ModuleExportModel
    from oracle import query as oq

    print(f"[composite] module export vs oracle ({mode})", flush=True)
    m = _make("quad", torch.float16, device, seed=9)
    cfg = SimpleNamespace(mode="dynamic", geometry=SimpleNamespace(max_distance_traveled=57.3))
    em = ModuleExportModel(config=cfg, model=m, maximum=2.4, minimum=0.15, dtype=torch.float16)
    axes = [oq.linspace_device(0.05, 0.95, n) for n in (9, 10, 11)]
    t = 0.3
    raw = em.query([torch.from_numpy(a) for a in axes], t, oq.RESCALE[mode], identity=True)
    got = em.query([torch.from_numpy(a) for a in axes], t, oq.RESCALE[mode])
    raw = raw.flatten().cpu().numpy()
    ref_rescaled = oq.rescale(raw, 57.3, 2.4 - 0.15, 0.15, mode)
    assert np.array_equal(got.flatten().cpu().numpy(), ref_rescaled)
    pts = oq.grid_points(axes)
    p16 = m.net.params.detach().to(torch.float16).double().cpu()
    with torch.no_grad():
        ref = on.composite_forward(_oracle_parts("quad", True), _oracle_cfg(m.net), p16, on.quad_inputs(pts, t),
                                   round_dtype=torch.float16).numpy()
    err = np.abs(raw - ref).max() / np.abs(ref).max()
    assert err < YTOL[torch.float16], err


def test_module_accumulated_step_with_one_subangle_equals_step(device):
    """The module path's continuous-scanning step with K = 1 (grouped schedule) gives the plain step's line
    integrals and gradient (float addition order aside)."""
    from nect_amd.module_step import ModuleTrainStep
    from nect_amd.sampling.geometry import Geometry

    print("[composite] module accumulated step, K = 1", flush=True)
    m = _make("quad", torch.float32, device, seed=11)
    geo = Geometry(nDetector=(64, 64), dDetector=(1.0, 1.0), mode="cone", DSD=1000.0, DSO=500.0,
                   nVoxel=(64, 64, 64), dVoxel=(0.5, 0.5, 0.5), angles=[0.9])
    cg = geo.get_c_geometry()
    S, B, t = 64, 2048, 0.2
    gen = torch.Generator(device=device).manual_seed(4)
    idx = torch.randperm(64 * 64, device=device, generator=gen)
    jit = torch.rand(B, device=device, generator=gen) - 0.5
    off = torch.zeros((B, 2), device=device)
    tgt = torch.rand(B, device=device, generator=gen) * 0.3
    l_max = geo.max_distance_traveled
    outs = []
    for acc in (False, True):
        st = ModuleTrainStep(m, torch.float32, B, S, loss="L2", dynamic=True)
        st.GROUPS_MIN_WORKGROUPS = 0
        if acc:
            st.step_accumulated(cg, [0.9], idx, 0, B, S, l_max / S, True, [jit], [off], tgt, l_max, lr=0.0, t=t,
                                optimizer_step=False)
        else:
            st.step(cg, 0.9, idx, 0, B, S, l_max / S, True, jit, off, tgt, l_max, lr=0.0, t=t, optimizer_step=False)
        torch.cuda.synchronize()
        outs.append((st.y_pred[:B].clone(), st.grad.clone(), float(st.loss_value)))
    (y0, g0, l0), (y1, g1, l1) = outs
    assert torch.allclose(y0, y1, rtol=1e-6, atol=1e-7)
    assert abs(l0 - l1) <= 1e-6 * abs(l0)
    assert float((g0 - g1).abs().max() / g0.abs().max()) < 1e-5
