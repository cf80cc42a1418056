"""GPU tests of the tiny-cuda-nn drop-in modules (autograd through the HIP kernels) and of the
reconstruct/trainer entry point end to end."""
import math

import numpy as np
import pytest
import torch

from oracle import hashgrid as hg
from oracle.train import model_forward

pytestmark = pytest.mark.gpu

ENC = {"otype": "HashGrid", "n_levels": 16, "n_features_per_level": 2, "log2_hashmap_size": 17,
       "base_resolution": 16, "per_level_scale": 1.5}
NET = {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "Softplus", "n_neurons": 64,
       "n_hidden_layers": 2}


@pytest.mark.parametrize("dtype,hidden", [(torch.float32, 2), (torch.float16, 2), (torch.float32, 6)])
def test_network_with_encoding_forward_backward_match_oracle(device, dtype, hidden):
    """NetworkWithInputEncoding forward and the parameter gradient of sum(w*y) against the oracle model (float32
    with 6 hidden layers: past the float32 fused backward, so the GEMM network)."""
    from nect_amd import tcnn

    print(f"[modules] network with encoding {dtype} {hidden} hidden layers", flush=True)
    model = tcnn.NetworkWithInputEncoding(3, 1, ENC, {**NET, "n_hidden_layers": hidden}, dtype=dtype).to(device)
    g = torch.Generator().manual_seed(3)
    n_net = model.desc.n_params
    with torch.no_grad():
        model.params[n_net:] = ((torch.rand(model.n_enc_params, generator=g) * 2 - 1) * 0.5).to(device)
        model.params.copy_(model.params.to(dtype).float())  # the oracle sees the values the kernels see
    x = torch.rand(3000, 3, generator=g)
    w = torch.randn(3000, generator=g).double()
    y = model(x.to(device))
    (y[:, 0].double() * w.to(device)).sum().backward()
    torch.cuda.synchronize()
    spec = hg.GridSpec(3, 16, 2, 17, 16, 1.5)
    cfg = {"n_params": n_net, "n_input_padded": 32, "n_hidden_layers": hidden, "n_output_padded": 16,
           "activation": "ReLU", "output_activation": "Softplus"}
    p = model.params.detach().double().cpu().requires_grad_(True)
    ref = model_forward(spec, cfg, p, x.numpy(), round_dtype=None if dtype == torch.float32 else dtype)
    (ref * w).sum().backward()
    ey = (y[:, 0].double().cpu() - ref.detach()).abs().max() / ref.detach().abs().max()
    eg = (model.params.grad.double().cpu() - p.grad).abs().max() / p.grad.abs().max()
    tol_y, tol_g = (1e-5, 2e-4) if dtype == torch.float32 else (3e-3, 3e-2)
    assert ey < tol_y, float(ey)
    assert eg < tol_g, float(eg)


def test_encoding_and_network_modules_compose(device):
    """Encoding followed by Network equals NetworkWithInputEncoding on the same parameters."""
    from nect_amd import tcnn

    enc = tcnn.Encoding(3, ENC, dtype=torch.float16).to(device)
    net = tcnn.Network(enc.n_output_dims, 1, NET, dtype=torch.float16).to(device)
    both = tcnn.NetworkWithInputEncoding(3, 1, ENC, NET, dtype=torch.float16).to(device)
    with torch.no_grad():
        both.params.copy_(torch.cat([net.params, enc.params]))
    x = torch.rand(777, 3, device=device)
    a = net(enc(x))
    b = both(x)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


# the static preset's encoder and network (nect/cfg/static/hash_grid.yaml) with a 2^14 table and a Softplus
# output; config_override replaces whole entries (the reference's cfg.update)
POOR_ENCODER_T14 = {"otype": "HashGrid", "n_levels": 19, "n_features_per_level": 4, "log2_hashmap_size": 14,
                    "base_resolution": 8, "max_resolution_factor": 2}
PRESET_NET_SOFTPLUS = {"otype": "FullyFusedMLP", "activation": "LeakyReLU", "output_activation": "Softplus",
                       "n_neurons": 64, "n_hidden_layers": 6}


@pytest.mark.parametrize("output_activation", ["None", "Softplus"])
def test_wide_network_state_is_freed_without_the_garbage_collector(device, output_activation):
    """A GEMM network's forward state (every layer's activations) is released when the backward is done and the
    output dropped, with Python's cycle collector off: nothing the autograd node keeps refers back to the
    output (an identity output activation returns the last GEMM's result itself)."""
    import gc

    from nect_amd.tcnn import Network

    cfg = {"otype": "CutlassMLP", "activation": "ReLU", "output_activation": output_activation, "n_neurons": 128,
           "n_hidden_layers": 2}
    net = Network(32, 1, cfg, dtype=torch.float16).to(device)
    x = torch.rand(65536, 32, device=device)
    gc.collect()
    gc.disable()
    try:
        net(x).sum().backward()          # params.grad now exists and is accumulated in place below
        torch.cuda.synchronize()
        base = torch.cuda.memory_allocated(device)
        for _ in range(3):
            net(x).sum().backward()
        torch.cuda.synchronize()
        grown = torch.cuda.memory_allocated(device) - base
    finally:
        gc.enable()
    # three steps' activations would be 3 x 65536 x (2 x 128 x 2 + 4) bytes = 101 MB
    assert grown < (1 << 20), grown


def test_reconstruct_recovers_a_sphere(device):
    """nect_amd.reconstruct (static, poor preset) on analytic projections of one sphere: the volume is finite,
    and brighter inside the sphere than outside.  On this tiny case (256 rays x 100 samples per step) the
    preset's learning rate, sized for 5 M points per batch (config.py:398-399), makes late training unstable:
    its ReLU output unit can die (the output is then 0 everywhere) or the loss can blow up, depending on the
    order of the float atomic additions.  The test therefore sets lr = 1e-3 with a 50-step warm-up, a Softplus
    output and 20 epochs (the rest of the preset as it is)."""
    from nect_amd import phantom
    from nect_amd.reconstruct import reconstruct
    from nect_amd.sampling.geometry import Geometry

    print("[modules] reconstruct sphere", flush=True)
    n_views = 24
    angles = [2 * math.pi * i / n_views for i in range(n_views)]
    geo = Geometry(nDetector=(32, 32), dDetector=(1.0, 1.0), mode="cone", DSD=400.0, DSO=200.0,
                   nVoxel=(32, 32, 32), dVoxel=(0.5, 0.5, 0.5), angles=angles)
    ph = phantom.Phantom(torch.tensor([[0.0, 0.0, 0.0]], dtype=torch.float64),
                         torch.tensor([[4.0, 4.0, 4.0]], dtype=torch.float64),
                         torch.eye(3, dtype=torch.float64)[None], torch.tensor([1.0], dtype=torch.float64))
    proj = phantom.project(ph, "cone", (32, 32), (1.0, 1.0), 200.0, 400.0, angles, device=device).cpu()
    vol = reconstruct(geo, proj / proj.max(), quality="poor", niter=20, verbose=False,
                      config_override={"encoder": POOR_ENCODER_T14, "net": PRESET_NET_SOFTPLUS,
                                       "lr": 1e-3, "warmup": {"steps": 50, "lr0": 0.01, "otype": "exponential"}})
    assert vol.shape == (32, 32, 32)
    assert np.all(np.isfinite(vol))
    c = np.arange(32) - 15.5
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(zz ** 2 + yy ** 2 + xx ** 2) * 0.5
    inside = vol[r < 3.0].mean()
    outside = vol[r > 6.0].mean()
    assert inside > 3 * abs(outside) and inside > 0, (inside, outside)


def test_poor_preset_as_shipped_stays_finite(device):
    """The poor preset with its own network (ReLU output) and learning rate, only the table shrunk to 2^14, on
    the sphere case: the 16-bit step skips any step whose gradient is not finite, so parameters and volume stay
    finite (whether the single ReLU output unit survives the preset's learning rate on a batch this small is
    the preset's business, not checked)."""
    from nect_amd import phantom
    from nect_amd.reconstruct import reconstruct
    from nect_amd.sampling.geometry import Geometry

    print("[modules] poor preset as shipped", flush=True)
    n_views = 24
    angles = [2 * math.pi * i / n_views for i in range(n_views)]
    geo = Geometry(nDetector=(32, 32), dDetector=(1.0, 1.0), mode="cone", DSD=400.0, DSO=200.0,
                   nVoxel=(32, 32, 32), dVoxel=(0.5, 0.5, 0.5), angles=angles)
    ph = phantom.Phantom(torch.tensor([[0.0, 0.0, 0.0]], dtype=torch.float64),
                         torch.tensor([[4.0, 4.0, 4.0]], dtype=torch.float64),
                         torch.eye(3, dtype=torch.float64)[None], torch.tensor([1.0], dtype=torch.float64))
    proj = phantom.project(ph, "cone", (32, 32), (1.0, 1.0), 200.0, 400.0, angles, device=device).cpu()
    vol = reconstruct(geo, proj / proj.max(), quality="poor", niter=20, verbose=False,
                      config_override={"encoder": POOR_ENCODER_T14})
    assert vol.shape == (32, 32, 32)
    assert np.all(np.isfinite(vol))


def test_trainer_continuous_scanning_runs(device):
    """The trainer with continous_scanning / accumulation_steps = 3 trains (loss decreases) on a small case."""
    from nect_amd import phantom
    from nect_amd.config import get_static_cfg, setup_cfg
    from nect_amd.sampling.geometry import Geometry
    from nect_amd.trainer import Trainer

    print("[modules] continuous-scanning trainer", flush=True)
    n_views = 12
    angles = [2 * math.pi * i / n_views for i in range(n_views)]
    geo = Geometry(nDetector=(32, 32), dDetector=(1.0, 1.0), mode="cone", DSD=400.0, DSO=200.0,
                   nVoxel=(32, 32, 32), dVoxel=(0.5, 0.5, 0.5), angles=angles)
    ph = phantom.ellipsoids((16.0, 16.0, 16.0), n=4, seed=2)
    proj = phantom.project(ph, "cone", (32, 32), (1.0, 1.0), 200.0, 400.0, angles, device=device).cpu()
    cfg = get_static_cfg("hash_grid")
    cfg.update({"epochs": 6, "continous_scanning": True, "accumulation_steps": 3, "loss": "L2", "lr": 1e-2,
                "encoder": {"otype": "HashGrid", "n_levels": 16, "n_features_per_level": 2, "log2_hashmap_size": 14,
                            "base_resolution": 8, "per_level_scale": 1.3},
                "net": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "Softplus",
                        "n_neurons": 64, "n_hidden_layers": 2},
                "geometry": geo})
    cfg["downsampling_detector"] = {"start": 1, "end": 1, "update_interval": 1000}
    cfg["points_per_ray"] = {"start": 32, "end": 32, "update_interval": 1000, "linear": True}
    # the default 500-step exponential warm-up would keep the learning rate near 1e-5 for this short run
    cfg["warmup"] = {"steps": 4, "lr0": 0.1, "otype": "exponential"}
    trainer = Trainer(setup_cfg(cfg), proj / proj.max(), save_ckpt=False, save_last=False, verbose=False)
    losses = []
    fused_step = trainer.fused.step_accumulated

    def record(*args, **kw):
        out = fused_step(*args, **kw)
        losses.append(float(out.item()))
        return out

    trainer.fused.step_accumulated = record
    trainer.fit()
    assert len(losses) == 6 * n_views
    assert np.isfinite(losses).all()
    assert np.mean(losses[-n_views:]) < 0.5 * np.mean(losses[:n_views]), (losses[:3], losses[-3:])


def test_trainer_checkpoint_round_trip(device, tmp_path):
    """A trainer's checkpoint (the reference's layout: net.params + torch Adam state) resumes a new trainer
    with the same parameters, Adam moments, position and learning rate; a checkpoint without this trainer's
    schedule entries (as the reference writes them) resumes with the schedule replayed to the same rate."""
    from nect_amd import phantom
    from nect_amd.config import get_static_cfg, setup_cfg
    from nect_amd.sampling.geometry import Geometry
    from nect_amd.trainer import Trainer

    print("[modules] trainer checkpoint round trip", flush=True)
    n_views = 8
    angles = [2 * math.pi * i / n_views for i in range(n_views)]
    geo = Geometry(nDetector=(32, 32), dDetector=(1.0, 1.0), mode="cone", DSD=400.0, DSO=200.0,
                   nVoxel=(32, 32, 32), dVoxel=(0.5, 0.5, 0.5), angles=angles)
    ph = phantom.ellipsoids((16.0, 16.0, 16.0), n=3, seed=4)
    proj = (phantom.project(ph, "cone", (32, 32), (1.0, 1.0), 200.0, 400.0, angles, device=device).cpu())
    cfg = get_static_cfg("hash_grid")
    cfg.update({"epochs": 2, "loss": "L2", "lr": 1e-3,
                "encoder": {"otype": "HashGrid", "n_levels": 16, "n_features_per_level": 2, "log2_hashmap_size": 14,
                            "base_resolution": 8, "per_level_scale": 1.3},
                "net": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "Softplus",
                        "n_neurons": 64, "n_hidden_layers": 2},
                "geometry": geo, "warmup": {"steps": 5, "lr0": 0.1, "otype": "exponential"},
                "points_per_ray": {"start": 32, "end": 32, "update_interval": 1000, "linear": True},
                "downsampling_detector": {"start": 1, "end": 1, "update_interval": 1000}})
    a = Trainer(setup_cfg(cfg), proj / proj.max(), output_directory=tmp_path / "a", save_ckpt=True, save_last=True,
                verbose=False, prune=False)
    a.fit()
    ck = tmp_path / "a" / "checkpoints" / "last.ckpt"
    assert ck.exists()
    st = torch.load(ck, map_location="cpu", weights_only=True)
    assert set(st["model"]) == {"net.params"} and set(st["optim"]["state"][0]) == {"step", "exp_avg", "exp_avg_sq"}
    b = Trainer(setup_cfg(cfg), proj / proj.max(), checkpoint=ck, save_ckpt=False, save_last=False, verbose=False)
    assert torch.equal(b.fused.params, a.fused.params)
    assert torch.equal(b.fused.m, a.fused.m) and torch.equal(b.fused.v, a.fused.v)
    assert b.fused.adam_steps() == a.fused.adam_steps() > 0
    assert (b.current_epoch, b.current_angle) == (a.current_epoch, a.current_angle)
    lr_a = a._lr_holder.param_groups[0]["lr"]
    assert b._lr_holder.param_groups[0]["lr"] == pytest.approx(lr_a, rel=1e-9)
    ref_like = {k: v for k, v in st.items() if k not in ("lr_warmup", "projection", "lr_scheduler")}
    torch.save(ref_like, tmp_path / "ref_like.ckpt")
    c = Trainer(setup_cfg(cfg), proj / proj.max(), checkpoint=tmp_path / "ref_like.ckpt", save_ckpt=False,
                save_last=False, verbose=False)
    assert torch.equal(c.fused.params, a.fused.params)
    assert c._lr_holder.param_groups[0]["lr"] == pytest.approx(lr_a, rel=1e-6)


def test_trainer_resumes_mid_run_on_the_schedule(device, tmp_path):
    """Resuming from a checkpoint taken mid-run (epoch 1, angle 3) with a changing detector binning and
    points-per-ray schedule continues with the views, binning and sample counts of the uninterrupted run
    (base_trainer.py:199-204: the schedule re-derived from the projection number; the view order of each
    epoch replayed)."""
    from nect_amd import phantom
    from nect_amd.config import get_static_cfg, setup_cfg
    from nect_amd.sampling.geometry import Geometry
    from nect_amd.trainer import Trainer

    print("[modules] trainer resume mid-run", flush=True)
    n_views = 8
    angles = [2 * math.pi * i / n_views for i in range(n_views)]
    geo = Geometry(nDetector=(32, 32), dDetector=(1.0, 1.0), mode="cone", DSD=400.0, DSO=200.0,
                   nVoxel=(32, 32, 32), dVoxel=(0.5, 0.5, 0.5), angles=angles)
    ph = phantom.ellipsoids((16.0, 16.0, 16.0), n=3, seed=5)
    proj = phantom.project(ph, "cone", (32, 32), (1.0, 1.0), 200.0, 400.0, angles, device=device).cpu()
    cfg = get_static_cfg("hash_grid")
    cfg.update({"epochs": 2, "loss": "L2", "lr": 1e-3,
                "encoder": {"otype": "HashGrid", "n_levels": 16, "n_features_per_level": 2, "log2_hashmap_size": 14,
                            "base_resolution": 8, "per_level_scale": 1.3},
                "net": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "Softplus",
                        "n_neurons": 64, "n_hidden_layers": 2},
                "geometry": geo, "warmup": {"steps": 5, "lr0": 0.1, "otype": "exponential"},
                "points_per_ray": {"start": 16, "end": 24, "update_interval": 2, "linear": True},
                "downsampling_detector": {"start": 4, "end": 1, "update_interval": 3}})

    def record(tr):
        seen = []
        step = tr.fused.step

        def wrapped(geometry, angle, *args, **kw):
            seen.append((tr.current_epoch, round(float(angle), 9), tr.downsample, tr.points_per_ray))
            return step(geometry, angle, *args, **kw)

        tr.fused.step = wrapped
        return seen

    a = Trainer(setup_cfg(cfg), proj / proj.max(), output_directory=tmp_path / "a", save_ckpt=True, save_last=True,
                verbose=False, prune=False)
    seen_a = record(a)
    a.fit()
    assert len(seen_a) == 2 * n_views
    assert len({s[2] for s in seen_a}) > 1 and len({s[3] for s in seen_a}) > 1
    st = a.state_dict()
    st.update({"epoch": 1, "angle": 3, "projection": n_views + 3})
    torch.save(st, tmp_path / "mid.ckpt")
    b = Trainer(setup_cfg(cfg), proj / proj.max(), checkpoint=tmp_path / "mid.ckpt", save_ckpt=False,
                save_last=False, verbose=False)
    seen_b = record(b)
    b.fit()
    assert seen_b == seen_a[n_views + 3:], (seen_b[:3], seen_a[n_views + 3:n_views + 6])


def test_reconstruct_from_projection_files(device, tmp_path):
    """reconstruct() given a directory of one-projection .npy files (the reference's file layout) trains from
    them: the volume is finite and brighter inside the sphere than outside."""
    from nect_amd import phantom
    from nect_amd.reconstruct import reconstruct
    from nect_amd.sampling.geometry import Geometry

    print("[modules] reconstruct from files", flush=True)
    n_views = 24
    angles = [2 * math.pi * i / n_views for i in range(n_views)]
    geo = Geometry(nDetector=(32, 32), dDetector=(1.0, 1.0), mode="cone", DSD=400.0, DSO=200.0,
                   nVoxel=(32, 32, 32), dVoxel=(0.5, 0.5, 0.5), angles=angles)
    ph = phantom.Phantom(torch.tensor([[0.0, 0.0, 0.0]], dtype=torch.float64),
                         torch.tensor([[4.0, 4.0, 4.0]], dtype=torch.float64),
                         torch.eye(3, dtype=torch.float64)[None], torch.tensor([1.0], dtype=torch.float64))
    proj = phantom.project(ph, "cone", (32, 32), (1.0, 1.0), 200.0, 400.0, angles, device=device).cpu().numpy()
    d = tmp_path / "projections"
    d.mkdir()
    for i in range(n_views):
        np.save(d / f"proj_{i:04d}.npy", proj[i].astype(np.float32))
    vol = reconstruct(geo, d, quality="poor", niter=20, verbose=False,
                      config_override={"encoder": POOR_ENCODER_T14, "net": PRESET_NET_SOFTPLUS,
                                       "lr": 1e-3, "warmup": {"steps": 50, "lr0": 0.01, "otype": "exponential"}})
    assert (d / "_min_max.yaml").exists()
    assert np.all(np.isfinite(vol))
    c = np.arange(32) - 15.5
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(zz ** 2 + yy ** 2 + xx ** 2) * 0.5
    inside, outside = vol[r < 3.0].mean(), vol[r > 6.0].mean()
    assert inside > 3 * abs(outside) and inside > 0, (inside, outside)
