"""CPU checks of the host-side configuration mirror (nect/config.py formulas) and geometry."""
import math

import pytest
import torch

from nect_amd.config import get_static_cfg, setup_cfg
from nect_amd.sampling.geometry import Geometry


def geom(**kw):
    base = dict(nDetector=(64, 64), dDetector=(1.0, 1.0), mode="cone", DSD=1000.0, DSO=500.0, nVoxel=(64, 64, 64),
                dVoxel=(0.5, 0.5, 0.5), angles=[i * 2 * math.pi / 90 for i in range(90)])
    base.update(kw)
    return Geometry(**base)


def test_geometry_lmax_and_c_struct():
    g = geom()
    ml = math.hypot(31.5, 31.5)
    assert abs(g.max_distance_traveled - ml / math.cos(math.atan(15.75 / (500 + ml / 2)))) < 1e-12
    g.set_detector_binning(2)
    c = g.get_c_geometry()
    assert c.nDetector == (32, 32) and c.dDetector == (2.0, 2.0)
    assert c.distance_origin_to_detector == 500.0
    with pytest.raises(ValueError):
        Geometry(nDetector=(4, 4), dDetector=(1.0, 1.0), mode="fan")
    with pytest.raises(ValueError):
        Geometry(nDetector=(4, 4.5), dDetector=(1.0, 1.0), mode="parallel", nVoxel=(4, 4, 4), dVoxel=(1, 1, 1))


def test_per_level_scale_formula():
    cfg = get_static_cfg("hash_grid")
    cfg["geometry"] = geom()
    c = setup_cfg(cfg)
    e = cfg["encoder"]
    expect = (e["max_resolution_factor"] * 64 / e["base_resolution"]) ** (1 / (e["n_levels"] - 1))
    assert abs(c.per_level_scale() - expect) < 1e-15
    cfg["encoder"]["per_level_scale"] = 1.5
    assert setup_cfg(cfg).per_level_scale() == 1.5


def test_lr_epochs_and_loss():
    cfg = get_static_cfg("hash_grid")
    cfg["geometry"] = geom()
    cfg["points_per_batch"] = 2_000_000
    c = setup_cfg(cfg)
    assert abs(c.learning_rate() - 0.001 * math.sqrt(2.0)) < 1e-15
    assert c.resolve_epochs(90) == math.ceil(49 / 90 * 64)
    cfg2 = get_static_cfg("hash_grid")
    cfg2.update(geometry=geom(), epochs="0.3x", loss="L1+L2")
    c2 = setup_cfg(cfg2)
    assert c2.resolve_epochs(90) == math.ceil(math.ceil(49 / 90 * 64) * 0.3)
    assert c2.loss_name() == "L1"  # "l1" is tested first, as in the reference


def test_schedules_follow_torch_schedulers():
    cfg = get_static_cfg("hash_grid")
    cfg.update(geometry=geom(), epochs=3, lr=0.01)
    c = setup_cfg(cfg)
    holder = torch.optim.SGD([torch.zeros(1)], lr=0.01)
    warm, decay = c.schedulers(holder, 90)
    lrs = []
    for p in range(1, 300):
        lrs.append(holder.param_groups[0]["lr"])
        decay.step()
        if p <= 500:
            warm.step()
    # exponential warm-up from lr0 * lr, monotone over the first steps
    assert abs(lrs[0] - 0.01 * (1 / ((1 / 0.001) ** (1 / 500)) ** 500)) < 1e-9
    assert all(b > a for a, b in zip(lrs[:50], lrs[1:51]))


def test_checkpoint_layout_reads_a_torch_adam_checkpoint(tmp_path):
    """A checkpoint in the reference's layout (base_trainer.py:452-476: the HashGrid module's state dict, whose
    one tensor is net.params, and torch.optim.Adam's state dict) is read back by read_checkpoint_state, and
    checkpoint_state writes that layout: torch's Adam loads what it writes."""
    import torch
    from torch import nn

    from nect_amd.trainer import checkpoint_state, read_checkpoint_state

    class Net(nn.Module):
        def __init__(self):
            super().__init__()
            self.params = nn.Parameter(torch.randn(1000, generator=torch.Generator().manual_seed(0)))

    class HashGridLike(nn.Module):
        def __init__(self):
            super().__init__()
            self.net = Net()

    model = HashGridLike()
    opt = torch.optim.Adam(model.parameters(), lr=1e-3, betas=(0.9, 0.99), eps=1e-15)
    for _ in range(3):
        opt.zero_grad()
        (model.net.params ** 2).sum().backward()
        opt.step()
    path = tmp_path / "ref.ckpt"
    torch.save({"model": model.state_dict(), "optim": opt.state_dict(), "epoch": 2, "angle": 5}, path)
    got = read_checkpoint_state(torch.load(path, map_location="cpu", weights_only=True))
    assert torch.equal(got["params"], model.net.params.detach())
    assert torch.equal(got["m"], opt.state[model.net.params]["exp_avg"])
    assert torch.equal(got["v"], opt.state[model.net.params]["exp_avg_sq"])
    assert got["step"] == 3
    # and the other way: what checkpoint_state writes loads into a torch module and a torch Adam
    st = checkpoint_state(got["params"] * 2, got["m"], got["v"], 7, 1e-3, (0.9, 0.99), 1e-15)
    torch.save(st, tmp_path / "ours.ckpt")
    st = torch.load(tmp_path / "ours.ckpt", map_location="cpu", weights_only=True)
    model2 = HashGridLike()
    model2.load_state_dict(st["model"])
    opt2 = torch.optim.Adam(model2.parameters(), lr=1e-3)
    opt2.load_state_dict(st["optim"])
    assert torch.equal(model2.net.params.detach(), got["params"] * 2)
    assert torch.equal(opt2.state[model2.net.params]["exp_avg"], got["m"])
    assert int(opt2.state[model2.net.params]["step"]) == 7
    assert read_checkpoint_state(checkpoint_state(got["params"], None, None, 0, 1e-3))["m"] is None


def test_load_projections_files_and_directories(tmp_path):
    """load_projections (nect/data.py:99-197): one .npy/.raw stack scaled by its own min/max, a directory of
    one-projection files scaled by the global min/max (and _min_max.yaml written, then read back), channel
    order WH/NWH transposed to HW, flip reversing both detector axes."""
    import numpy as np
    import yaml

    from nect_amd.trainer import load_projections

    rng = np.random.default_rng(5)
    stack = (rng.random((3, 4, 6)) * 7 + 2).astype(np.float32)        # N=3, H=4, W=6
    want = (stack - stack.min()) / (stack - stack.min()).max()
    np.save(tmp_path / "p.npy", stack)
    assert np.allclose(load_projections(tmp_path / "p.npy").numpy(), want, rtol=1e-6)
    stack.tofile(tmp_path / "p.raw")
    assert np.allclose(load_projections(tmp_path / "p.raw", (4, 6)).numpy(), want, rtol=1e-6)
    # NWH stack: stored transposed, read back as NHW
    np.save(tmp_path / "t.npy", np.ascontiguousarray(stack.transpose(0, 2, 1)))
    assert np.allclose(load_projections(tmp_path / "t.npy", channel_order="NWH").numpy(), want, rtol=1e-6)
    assert np.allclose(load_projections(tmp_path / "p.npy", flip=True).numpy(), want[:, ::-1, ::-1], rtol=1e-6)
    # a directory of one-projection files, mixed .npy and .raw, sorted by name
    d = tmp_path / "dir"
    d.mkdir()
    np.save(d / "a0.npy", stack[0])
    stack[1].tofile(d / "a1.raw")
    np.save(d / "a2.npy", stack[2])
    lo, hi = float(stack.min()), float(stack.max())
    got = load_projections(d, (4, 6)).numpy()
    assert np.allclose(got, (stack - lo) / (hi - lo), rtol=1e-6)
    lim = yaml.safe_load(open(d / "_min_max.yaml"))
    assert lim["min"] == lo and lim["max"] == hi
    # the cached range wins over the files' own (the reference reads it back the same way)
    yaml.safe_dump({"min": 0.0, "max": 20.0}, open(d / "_min_max.yaml", "w"))
    assert np.allclose(load_projections(d, (4, 6)).numpy(), stack / 20.0, rtol=1e-6)


def test_load_projections_reads_tiff_and_png_files(tmp_path):
    """One-projection image files (nect/data.py:169-176): float32 TIFF and 16-bit PNG read with PIL, in a
    directory walked in sorted order and scaled by the global min/max like .npy files."""
    import numpy as np
    from PIL import Image

    from nect_amd.trainer import load_projections

    rng = np.random.default_rng(9)
    f32 = (rng.random((2, 5, 7)) * 3 + 1).astype(np.float32)
    d = tmp_path / "tif"
    d.mkdir()
    for i in range(2):
        Image.fromarray(f32[i]).save(d / f"p{i}.tif")
    lo, hi = float(f32.min()), float(f32.max())
    assert np.allclose(load_projections(d).numpy(), (f32 - lo) / (hi - lo), rtol=1e-6)
    u16 = rng.integers(100, 60000, (2, 5, 7)).astype(np.uint16)
    d2 = tmp_path / "png"
    d2.mkdir()
    for i in range(2):
        Image.fromarray(u16[i]).save(d2 / f"p{i}.png")
    lo, hi = float(u16.min()), float(u16.max())
    assert np.allclose(load_projections(d2).numpy(), (u16.astype(np.float32) - lo) / (hi - lo), rtol=1e-6)
