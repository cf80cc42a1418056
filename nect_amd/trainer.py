"""Trainer: BaseTrainer (nect/trainers/base_trainer.py:36-666) on the fused HIP step.

Kept semantics:
  * seeding (L.seed_everything(42), base_trainer.py:63) and one random view
    order per epoch (DataLoader(shuffle=True), :78-80);
  * per view: detector binning and points-per-ray schedules (set_projection,
    :235-270), a fresh random pixel permutation (projector.py:72-97), targets
    gathered through it (projector.py:128-131), avg_pool2d binning of the
    projection (:307-313);
  * ``min(batch_per_proj, batch_per_epoch)`` batches per view with the
    projector's tail rule (projector.py:106-112);
  * learning rate from torch's LambdaLR warm-up and cosine/exponential decay,
    stepped per projection (:281-288), fed to the fused Adam;
  * create_volume on voxel centres, rescaled by L_max and the data range
    (:402-450); checkpoints with model, optimiser state, epoch and angle
    (:452-484).
What changes: the batch body is one fused step (nect_amd.fused) instead of
sampler -> mask/compact -> tcnn -> integral -> backward -> optimiser, and no
host synchronisation happens inside the loop.
"""
from __future__ import annotations

import datetime
import math
import os
import random
import time
from pathlib import Path

import numpy as np
import yaml
import torch
import torch.nn.functional as F

from . import _lib
from . import grid as _grid
from . import mlp as _mlp
from .config import Config
from .fused import FusedTrainStep
from .sampling import ct_sampling as cs
from .tcnn import NetworkWithInputEncoding


def seed_everything(seed: int = 42):
    random.seed(seed)
    np.random.seed(seed)
    torch.manual_seed(seed)


class ProjectionData:
    """In-memory projections (NeCTDatasetLoaded, nect/data.py:289-316): not rescaled, range recorded.

    ``data_range`` = (maximum, minimum) overrides the recorded range: projections read from files arrive
    scaled, and NeCTDataset then reports the range of the files (see load_projections)."""

    def __init__(self, projections, geometry, device="cpu", data_range=None):
        if isinstance(projections, np.ndarray):
            projections = torch.tensor(projections, dtype=torch.float32)
        if projections.dim() != 3:
            raise ValueError("Projections must have shape (N, H, W)")
        self.proj = projections.to(device=device, dtype=torch.float32)
        if data_range is None:
            self.maximum = self.proj.max()
            self.minimum = self.proj.min()
        else:
            self.maximum = torch.tensor(float(data_range[0]))
            self.minimum = torch.tensor(float(data_range[1]))
        n = self.proj.shape[0]
        if geometry.angles is None or len(geometry.angles) != n:
            raise ValueError("geometry.angles must have one angle per projection")
        self.angles = torch.tensor(np.asarray(geometry.angles), dtype=torch.float64)
        ts = geometry.timesteps
        if ts is not None:
            t = torch.tensor(np.asarray(ts), dtype=torch.float32)
            if t.numel() != n:
                raise ValueError("Number of timesteps must match number of images")
            self.timesteps = t / t.max() if t.max() > 1 else t       # nect/data.py:74-79
        else:
            self.timesteps = torch.linspace(0, 1, steps=n)           # nect/data.py:83

    def __len__(self):
        return self.proj.shape[0]

    def __getitem__(self, i):
        return self.proj[i], float(self.angles[i]), float(self.timesteps[i])


IMAGE_SUFFIXES = (".tif", ".tiff", ".png")
PROJECTION_SUFFIXES = (".npy", ".raw") + IMAGE_SUFFIXES


def _read_projection_file(path: Path, shape) -> np.ndarray:
    suffix = path.suffix.lower()
    if suffix == ".raw":
        return np.fromfile(path, dtype=np.float32).reshape(shape)      # nect/data.py:159-163
    if suffix in IMAGE_SUFFIXES:
        # nect/data.py:169-176: image files are read with PIL (one projection per file)
        from PIL import Image

        with Image.open(path) as im:
            return np.asarray(np.array(im), dtype=np.float32)
    return np.asarray(np.load(path, allow_pickle=False), dtype=np.float32)


def load_projections(path: str | Path, n_detector=None, channel_order: str | None = None,
                     flip: bool = False, return_range: bool = False):
    """Projections [N, H, W] from files, as NeCTDataset reads them (nect/data.py:99-197), scaled to [0, 1].

    ``path`` is one file holding every projection (channel order "nhw" by default, or "nwh") or a directory of
    one-projection files ("hw" by default, or "wh"), walked in sorted order; .npy, or .raw float32 (which
    needs ``n_detector`` = (rows, cols)).  One file is scaled by its own minimum and maximum; a directory by the
    global ones, kept in ``_min_max.yaml`` beside the files as the reference keeps them (read when present,
    written otherwise).  ``flip`` reverses both detector axes; "nwh"/"wh" files are transposed to (H, W).
    .tif/.tiff/.png projections (one per file) are read with PIL.

    With ``return_range`` the result is (projections, maximum, minimum), the range NeCTDataset reports and
    the exports rescale with: for one file the minimum and the maximum *after* subtracting it
    (nect/data.py:178-183), for a directory the global maximum and minimum (:132-143)."""
    path = Path(path)
    rows, cols = (int(n_detector[0]), int(n_detector[1])) if n_detector is not None else (-1, -1)
    if path.is_file():
        order = (channel_order or "nhw").lower()
        if order not in ("nhw", "nwh"):
            raise ValueError(f"channel order of a projection file must be NHW or NWH, not {channel_order}")
        dims = {"n": -1, "h": rows, "w": cols}
        if path.suffix.lower() == ".raw" and n_detector is None:
            raise ValueError(".raw projections need the detector shape")
        proj = _read_projection_file(path, [dims[c] for c in order])
        if proj.ndim != 3:
            raise ValueError(f"expected a stack of projections [N, H, W] in {path}, got shape {proj.shape}")
        proj = torch.from_numpy(proj.copy())
        minimum = proj.min()
        proj = proj - minimum                                            # nect/data.py:178-183
        maximum = proj.max()
        proj = proj / maximum
        rng = (float(maximum), float(minimum))
    elif path.is_dir():
        order = (channel_order or "hw").lower()
        if order not in ("hw", "wh"):
            raise ValueError(f"channel order of one-projection files must be HW or WH, not {channel_order}")
        files = []
        for root, _, names in sorted(os.walk(path)):
            for name in sorted(names):
                f = Path(root) / name
                if f.suffix.lower() in PROJECTION_SUFFIXES:
                    files.append(f)
        if not files:
            raise ValueError(f"no .npy/.raw projections in {path}")
        if any(f.suffix.lower() == ".raw" for f in files) and n_detector is None:
            raise ValueError(".raw projections need the detector shape")
        dims = {"h": rows, "w": cols}
        stack = np.stack([_read_projection_file(f, [dims[c] for c in order]) for f in files])
        mm = path / "_min_max.yaml"
        if mm.exists():                                                  # nect/data.py:108-114
            with open(mm) as fh:
                lim = yaml.safe_load(fh)
            lo, hi = float(lim["min"]), float(lim["max"])
        else:
            lo, hi = float(stack.min()), float(stack.max())
            with open(mm, "w") as fh:                                    # nect/data.py:134-136
                yaml.safe_dump({"max": hi, "min": lo}, fh)
        proj = (torch.from_numpy(stack) - lo) / (hi - lo)                # nect/data.py:140-143
        rng = (hi, lo)
    else:
        raise FileNotFoundError(path)
    if flip:
        proj = torch.flip(proj, [-2, -1])                                # nect/data.py:185-192
    if order in ("nwh", "wh"):
        proj = proj.transpose(-1, -2)                                    # nect/data.py:193-194
    if return_range:
        return proj.contiguous(), rng[0], rng[1]
    return proj.contiguous()


def data_range_of_config(config) -> tuple[float, float]:
    """(maximum, minimum) of a run's projections as NeCTDataset(config) reports it: the ``data_range`` this
    trainer stores in config.yaml, else recomputed from the projection files the config names (img_path)."""
    dr = config.raw.get("data_range")
    if dr is not None:
        return float(dr["maximum"]), float(dr["minimum"])
    img = config.raw.get("img_path")
    if img is None:
        raise ValueError("the run's config names no projection files (img_path) and holds no data_range")
    _, mx, mn = load_projections(img, config.geometry.nDetector, config.raw.get("channel_order"),
                                 bool(config.raw.get("flip", False)), return_range=True)
    return mx, mn


MODEL_KEY = "net.params"   # HashGrid.net (nect/network/networks.py:52) holds the flat tcnn parameter vector


def checkpoint_state(params: torch.Tensor, m: torch.Tensor | None, v: torch.Tensor | None, step: int, lr: float,
                     betas=(0.9, 0.999), eps: float = 1e-8, weight_decay: float = 0.0) -> dict:
    """Model and optimiser entries of a checkpoint in the reference's layout (base_trainer.py:452-476 saves
    {"model": HashGrid, "optim": torch.optim.Adam, ...} through Lightning Fabric, i.e. their state dicts): the
    model's only tensor is HashGrid.net.params, the flat float32 [MLP weights | grid table] vector this library
    trains; Adam's state is torch's (step, exp_avg, exp_avg_sq of parameter 0).  m = None leaves the optimiser
    out (save_optimizer=False)."""
    st = {"model": {MODEL_KEY: params.detach().float().cpu()}}
    if m is not None:
        st["optim"] = {
            "state": {0: {"step": torch.tensor(float(step)), "exp_avg": m.detach().float().cpu(),
                          "exp_avg_sq": v.detach().float().cpu()}},
            "param_groups": [{"lr": float(lr), "betas": (float(betas[0]), float(betas[1])), "eps": float(eps),
                              "weight_decay": float(weight_decay), "amsgrad": False, "maximize": False,
                              "foreach": None, "capturable": False, "differentiable": False, "fused": None,
                              "params": [0]}]}
    return st


def read_checkpoint_state(st: dict) -> dict:
    """params, m, v, step and lr from a checkpoint written by checkpoint_state or by the reference (the
    model's state dict holding one tcnn parameter vector, ``net.params``; Adam's state for parameter 0 and
    its learning rate).  m/v (lr) are None when the checkpoint holds no optimiser state."""
    model = st["model"]
    if MODEL_KEY in model:
        params = model[MODEL_KEY]
    else:
        tensors = [k for k, t in model.items() if torch.is_tensor(t)]
        if len(tensors) != 1:
            raise ValueError(f"expected one parameter vector in the model state, found {tensors}")
        params = model[tensors[0]]
    out = {"params": params.float(), "m": None, "v": None, "step": 0, "lr": None}
    opt = st.get("optim")
    if opt and opt.get("param_groups"):
        out["lr"] = float(opt["param_groups"][0]["lr"])
    if opt and opt.get("state"):
        s0 = opt["state"][0] if 0 in opt["state"] else next(iter(opt["state"].values()))
        out["m"], out["v"] = s0["exp_avg"].float(), s0["exp_avg_sq"].float()
        step = s0["step"]
        out["step"] = int(step.item() if torch.is_tensor(step) else step)
    return out


SUPPORTED_OPTIMIZERS = ("adam",)


class BaseTrainer:
    """BaseTrainer(config, output_directory, checkpoint, save_ckpt, save_last, save_optimizer, verbose, log,
    cancel_at, keep_two, prune) of nect/trainers/base_trainer.py:36-50: the projections are read from the files
    config["img_path"] names (nect/data.py:99-197) unless ``projections`` is given (ProjectionsLoadedTrainer).

    The run directory is ``output_directory`` itself (config.yaml, checkpoints/last.ckpt); the reference adds
    a config-name/date sub-directory below it.  With a checkpoint and no output directory the run directory is
    the checkpoint's (checkpoints/last.ckpt -> its grandparent).  ``dtype`` selects the compute precision."""

    def __init__(self, config: Config, output_directory: str | Path | None = None,
                 checkpoint: str | Path | None = None, save_ckpt: bool = True, save_last: bool = True,
                 save_optimizer: bool = True, verbose: bool = True, log: bool = False, cancel_at: str | None = None,
                 keep_two: bool = True, prune: bool = True, *, projections=None, data_range=None,
                 dtype: torch.dtype = torch.float16, device: str | torch.device = "cuda"):
        seed_everything(42)
        self.keep_two = keep_two
        self.prune = prune
        if checkpoint and output_directory is None:
            output_directory = Path(checkpoint).parent.parent
        if projections is None:
            img = config.raw.get("img_path")
            if img is None or not Path(str(img)).exists():
                raise ValueError(f"config img_path does not name projection files: {img}")
            projections, mx, mn = load_projections(img, config.geometry.nDetector, config.raw.get("channel_order"),
                                                   bool(config.raw.get("flip", False)), return_range=True)
            data_range = (mx, mn) if data_range is None else data_range
        # training options this trainer does not implement fail here rather than train something else
        otype = str(config["optimizer"].get("otype", "Adam")).casefold()
        if otype not in SUPPORTED_OPTIMIZERS:
            raise NotImplementedError(f"optimizer {config['optimizer'].get('otype')} is not implemented (Adam only)")
        for key in ("s3im", "add_poisson"):
            if config.raw.get(key):
                raise NotImplementedError(f"config option {key} is not implemented")
        self.config = config
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("the trainer runs on the GPU only")
        self.dtype = dtype
        self.verbose = verbose
        self.save = save_ckpt
        self.save_last = save_last
        self.save_optimizer = save_optimizer
        self.cancel_at = datetime.datetime.fromisoformat(cancel_at) if cancel_at else None
        self.geometry = config.geometry
        self.dataset = ProjectionData(projections, self.geometry, data_range=data_range)
        self.epochs = config.resolve_epochs(len(self.dataset))
        self.loss_name = config.loss_name()
        if (save_ckpt or save_last or log) and output_directory is None:
            raise ValueError("Output directory must be provided if logging or saving checkpoints")
        self.output_directory = Path(output_directory) if output_directory is not None else None

        # model: the module the config names (Config.get_model); a single hash grid with the 64-neuron network
        # trains on the fused step (its flat parameters trained in place), any other model on the module step
        from .mlp import MlpDesc
        from .networks import HashGrid, HashGrid4D, build_model

        self.model = build_model(config, dtype).to(self.device)
        net = self.model.net
        grids = net.grid_parts()
        self.fused_model = (isinstance(self.model, (HashGrid, HashGrid4D)) and len(net.parts) == 1
                            and len(grids) == 1 and isinstance(net.desc, MlpDesc))
        self.grid_desc = grids[0].desc if self.fused_model else None
        self.mlp_desc = net.desc if self.fused_model else None

        # schedules (base_trainer.py:132-175)
        self.downsample = int(config["downsampling_detector"]["start"])
        ppr = config["points_per_ray"]
        self.points_per_ray = int(ppr["start"])
        if ppr["end"] == "auto":
            ppr["end"] = math.ceil(max(self.geometry.nDetector) * 1.5)
        elif isinstance(ppr["end"], str):
            ppr["end"] = math.ceil(max(self.geometry.nDetector) * float(ppr["end"].split("x")[0]))
        if isinstance(ppr["update_interval"], str):
            factor = 0.9 if ppr["update_interval"] == "auto" else float(ppr["update_interval"].split("x")[0])
            total = len(self.dataset) * self.epochs * factor
            need = max(1, ppr["end"] - ppr["start"])
            ppr["update_interval"] = math.ceil(total / need)
        self.points_per_batch = config.points_per_batch() // 2          # base_trainer.py:143
        self.lr = config.learning_rate()
        opt = config["optimizer"]
        max_S = max(int(ppr["end"]), self.points_per_ray)
        max_rays = max(1, self.points_per_batch // self.points_per_ray)
        rays_cap = min(max_rays, self.geometry.nDetector[0] * self.geometry.nDetector[1])
        adam = dict(betas=(opt["beta1"], opt.get("beta2", 0.999)), eps=1e-8, weight_decay=opt["weight_decay"])
        if self.fused_model:
            self.fused = FusedTrainStep(self.grid_desc, self.mlp_desc, self.model.net.params.data, dtype,
                                        max_rays=rays_cap, max_points_per_ray=max_S, loss=self.loss_name, **adam)
        else:
            from .module_step import ModuleTrainStep

            self.fused = ModuleTrainStep(self.model, dtype, rays_cap, max_S, loss=self.loss_name,
                                         dynamic=config.mode == "dynamic", **adam)
        clip = config.raw.get("clip_grad_value")
        self.fused.clip_grad_value = None if clip is None else float(clip)
        # a stand-in optimiser only to drive torch's LR schedulers
        self._lr_holder = torch.optim.SGD([torch.zeros(1)], lr=self.lr)
        self.lr_warmup, self.lr_decay = config.schedulers(self._lr_holder, len(self.dataset))
        # continuous scanning (nect/config.py:269-270): the line integral averages accumulation_steps sub-angles
        self.accumulation_steps = int(config.raw.get("accumulation_steps") or 0) \
            if config.raw.get("continous_scanning") else 0
        if config.raw.get("continous_scanning") and self.accumulation_steps < 1:
            raise ValueError("accumulation_steps must be provided")
        self.batch_per_proj = config["batch_per_proj"]
        if self.batch_per_proj == "all":
            self.batch_per_proj = 1_000_000_000
        self.current_epoch = 0
        self.current_angle = 0
        self.current_projection = 0
        self.step = 0
        self.last_loss = None
        self._view_gen = torch.Generator().manual_seed(42)
        self._dev_gen = torch.Generator(device=self.device).manual_seed(42)
        if checkpoint:
            self.load(checkpoint)

    # ------------------------------------------------------------ projector

    def _set_projection(self, projection: int):
        """base_trainer.py:235-270 without the projector object."""
        c = self.config
        dd = c["downsampling_detector"]
        ppr = c["points_per_ray"]
        if projection != 0 and projection % dd["update_interval"] == 0 and self.downsample > dd["end"]:
            self.downsample //= 2
        if projection != 0 and projection % ppr["update_interval"] == 0 and self.points_per_ray < ppr["end"]:
            self.points_per_ray = self.points_per_ray + 1 if ppr.get("linear", True) else self.points_per_ray * 2
        self.geometry.set_detector_binning(self.downsample)
        self.c_geometry = self.geometry.get_c_geometry()
        self.random_offset_detector = 0.5 if self.downsample > 1 else 0.0
        self.total_pixels = (self.geometry.nDetector[0] // self.downsample) * (self.geometry.nDetector[1] // self.downsample)
        self.batch_size = self.points_per_batch // self.points_per_ray
        self.random_indexes = torch.randperm(self.total_pixels, dtype=torch.int64, device=self.device,
                                             generator=self._dev_gen)
        self.spacing = self.geometry.max_distance_traveled / self.points_per_ray
        self.batch_per_epoch = self.total_pixels // self.batch_size + (1 if self.total_pixels % self.batch_size else 0)

    def _projection(self, proj: torch.Tensor) -> torch.Tensor:
        proj = proj.to(self.device, non_blocking=True)
        if self.downsample != 1:
            proj = F.avg_pool2d(proj.unsqueeze(0), kernel_size=self.downsample, stride=self.downsample).squeeze(0)
        return proj.flatten()

    # ----------------------------------------------------------------- loop

    def fit(self):
        n = len(self.dataset)
        t0 = time.perf_counter()
        for epoch in range(self.current_epoch, self.epochs):
            order = torch.randperm(n, generator=self._view_gen).tolist()
            for i, view in enumerate(order):
                if i < self.current_angle:
                    continue
                proj, angle, t = self.dataset[view]
                self._set_projection(self.current_projection)
                proj = self._projection(proj)
                lr = self._lr_holder.param_groups[0]["lr"]
                for batch_num in range(min(int(self.batch_per_proj), self.batch_per_epoch)):
                    start = batch_num * self.batch_size
                    bs = self.batch_size
                    if (batch_num + 1) * bs > self.total_pixels:
                        if batch_num == 0:
                            bs = self.total_pixels % bs
                        else:
                            continue
                    target = proj[self.random_indexes[start:start + bs]].contiguous()
                    if self.accumulation_steps:
                        # continuous scanning (continous_scanning_trainer.py:81-167, data.py:221-238): the
                        # projection integrates over [angle_i, angle_{i+1}); midpoints of K equal parts
                        a0, a1 = self._angle_range(view)
                        edges = np.linspace(a0, a1, self.accumulation_steps + 1, endpoint=True)
                        angles = [float((edges[k] + edges[k + 1]) / 2) for k in range(self.accumulation_steps)]
                        jitters, offsets = [], []
                        for _ in angles:
                            jitters.append(torch.rand(bs, device=self.device, generator=self._dev_gen) - 0.5)
                            offsets.append((torch.rand((bs, 2), device=self.device, generator=self._dev_gen) - 0.5)
                                           * self.random_offset_detector)
                        self.last_loss = self.fused.step_accumulated(
                            self.c_geometry, angles, self.random_indexes, start, bs, self.points_per_ray,
                            self.spacing, self.config["uniform_ray_spacing"], jitters, offsets, target,
                            self.geometry.max_distance_traveled, lr=lr, t=t)
                    else:
                        jitter = torch.rand(bs, device=self.device, generator=self._dev_gen) - 0.5
                        offset = (torch.rand((bs, 2), device=self.device, generator=self._dev_gen) - 0.5) \
                            * self.random_offset_detector
                        self.last_loss = self.fused.step(self.c_geometry, angle, self.random_indexes, start, bs,
                                                         self.points_per_ray, self.spacing,
                                                         self.config["uniform_ray_spacing"], jitter, offset, target,
                                                         self.geometry.max_distance_traveled, lr=lr, t=t)
                    self.step += 1
                self.current_angle = i + 1
                self._on_angle_end()
            self.current_angle = 0
            self.current_epoch = epoch + 1
        torch.cuda.synchronize(self.device)
        self.training_time = time.perf_counter() - t0
        if self.save_last:
            self.save_model(last=True)

    def _angle_range(self, view: int) -> tuple[float, float]:
        """[angle_i, angle_{i+1}) of projection i; the last one extends by the previous spacing."""
        a = self.dataset.angles
        if view == len(a) - 1:
            return float(a[view]), float(a[view] + (a[view] - a[view - 1]))
        return float(a[view]), float(a[view + 1])

    def _on_angle_end(self):
        if self.cancel_at is not None and datetime.datetime.now(datetime.timezone.utc) > self.cancel_at:
            self.save_model(last=True)
            raise SystemExit(0)
        self.current_projection += 1
        self._lr_holder.step()  # no-op (no gradient); keeps torch's scheduler/optimizer ordering
        self.lr_decay.step()
        if self.current_projection <= self.config["warmup"]["steps"]:
            self.lr_warmup.step()

    # --------------------------------------------------------------- queries

    def evaluate(self, points: torch.Tensor, chunk: int = 1 << 22) -> torch.Tensor:
        """Model output mu [N] at points [N, D] (float32, in [0, 1]) from the current shadow parameters."""
        out = torch.empty(points.shape[0], dtype=torch.float32, device=self.device)
        F_ = self.grid_desc.n_output_dims
        for s in range(0, points.shape[0], chunk):
            x = points[s:s + chunk].contiguous()
            feats = _grid.encode(self.grid_desc, self.fused.table, x)
            _mlp.forward(self.mlp_desc, self.fused.mlp_weights, feats, out=out[s:s + chunk].view(-1, 1))
        return out

    def create_volume(self, save: bool = False, save_path: str | None = None, timestep: float | None = None,
                      cpu: bool = False) -> torch.Tensor:
        """Volume at the voxel centres (base_trainer.py:402-450) on the fused query kernel: coordinates
        linspace(0.5/n, 1 - 0.5/n, n) on the GPU (y and x over the grid padded by sample_outside, then cut back),
        values ((mu / L_max) * (max - min)) + min -- for a half-precision model rounded to half after each
        operation, as the reference computes it on the network's half output."""
        from .serving import RESCALE_RECIPROCAL, RESCALE_RECIPROCAL_F16, QueryRequest, QueryServer

        if self.config.mode == "dynamic" and timestep is None:
            return None
        if not self.fused_model:
            return self._module_volume(save, save_path, timestep, cpu)
        size = [int(v) for v in self.geometry.nVoxel]
        rm = int(self.config.get("sample_outside", 0) or 0)
        padded = [size[0], size[1] + 2 * rm, size[2] + 2 * rm]
        dev = self.device
        axes = [torch.linspace(0.5 / padded[k], 1 - 0.5 / padded[k], steps=padded[k], device=dev) for k in range(3)]
        if rm > 0:
            axes[1], axes[2] = axes[1][rm:-rm], axes[2][rm:-rm]
        axes = [a.contiguous() for a in axes]
        half = self.dtype == torch.float16
        r = QueryRequest(tuple(size), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0),
                         0.0 if timestep is None else float(timestep), l_max=self.geometry.max_distance_traveled,
                         span=(self.dataset.maximum - self.dataset.minimum).item(), lo=self.dataset.minimum.item(),
                         mu_half=half, rescale=RESCALE_RECIPROCAL_F16 if half else RESCALE_RECIPROCAL,
                         axes=tuple(axes))
        srv = QueryServer(self.grid_desc, self.mlp_desc, self.fused.mlp_weights, self.fused.table,
                          out_dtype=torch.float32, chunk=1 << 20, max_jobs=64)
        out = srv.run_alone(r)
        if cpu:
            out = out.cpu()
        if save:
            p = save_path or (self.output_directory / "final.npy")
            np.save(p, out.cpu().numpy())
        return out

    def _module_volume(self, save, save_path, timestep, cpu):
        """create_volume of a model the query kernel does not cover (composite encodings, wide networks): the
        reference's own slice loop (base_trainer.py:402-450) -- voxel centres by torch.linspace on the GPU, the
        network's output in the compute dtype (half, as tiny-cuda-nn returns it), then / L_max, * (max - min),
        + min as torch computes them on that tensor."""
        size = [int(v) for v in self.geometry.nVoxel]
        rm = int(self.config.get("sample_outside", 0) or 0)
        ss = [size[0], size[1] + 2 * rm, size[2] + 2 * rm]
        dev = self.device
        out = torch.zeros(size, dtype=torch.float32, device="cpu" if cpu else dev)
        z_lin = torch.linspace(0.5 / size[0], 1 - 0.5 / size[0], steps=size[0], device=dev)
        cut = slice(rm, -rm) if rm > 0 else slice(None)
        y_lin = torch.linspace(0.5 / ss[1], 1 - 0.5 / ss[1], steps=ss[1], device=dev)[cut]
        x_lin = torch.linspace(0.5 / ss[2], 1 - 0.5 / ss[2], steps=ss[2], device=dev)[cut]
        span = self.dataset.maximum.item() - self.dataset.minimum.item()
        with torch.no_grad():
            for i in range(size[0]):
                z, y, x = torch.meshgrid([z_lin[i], y_lin, x_lin], indexing="ij")
                grid = torch.stack((z.flatten(), y.flatten(), x.flatten())).t().contiguous()
                mu = self.model(grid, timestep) if self.config.mode == "dynamic" else self.model(grid)
                sl = mu.to(self.dtype).reshape(size[1], size[2])
                sl = sl / self.geometry.max_distance_traveled
                sl = sl * span
                sl = sl + self.dataset.minimum.item()
                out[i] = sl.cpu() if cpu else sl
        if save:
            p = save_path or (self.output_directory / "final.npy")
            np.save(p, out.cpu().numpy())
        return out

    # ----------------------------------------------------------- checkpoints

    def state_dict(self) -> dict:
        """The reference's checkpoint layout (model and Adam state; see checkpoint_state), plus the schedule."""
        lr = self._lr_holder.param_groups[0]["lr"]
        st = checkpoint_state(self.fused.params, self.fused.m if self.save_optimizer else None, self.fused.v,
                              self.fused.adam_steps(), lr, self.fused.betas, self.fused.eps, self.fused.weight_decay)
        st.update({"epoch": self.current_epoch, "angle": self.current_angle, "projection": self.current_projection,
                   "lr_scheduler": self.lr_decay.state_dict(), "lr_warmup": self.lr_warmup.state_dict()})
        return st

    def save_model(self, last: bool = False):
        """checkpoints/last.ckpt (the previous one kept as 2nd_last.ckpt with keep_two) and config.yaml
        (base_trainer.py:452-484); the final checkpoint of a pruned run holds no optimiser state."""
        if not (self.save or (last and self.save_last)) or self.output_directory is None:
            return
        ck = self.output_directory / "checkpoints"
        ck.mkdir(parents=True, exist_ok=True)
        last_p = ck / "last.ckpt"
        if last_p.exists() and self.keep_two:
            os.replace(last_p, ck / "2nd_last.ckpt")
        st = self.state_dict()
        if last and self.prune:
            st.pop("optim", None)
        torch.save(st, last_p)
        self.config.raw["data_range"] = {"maximum": float(self.dataset.maximum), "minimum": float(self.dataset.minimum)}
        self.config.to_yaml(self.output_directory / "config.yaml")

    def load(self, path):
        """Resumes from a checkpoint of this trainer or the model/Adam state of a reference checkpoint
        (base_trainer.py:189-195); a reference checkpoint's schedule is re-derived from epoch and angle."""
        st = torch.load(path, map_location="cpu", weights_only=True)
        got = read_checkpoint_state(st)
        if got["params"].numel() != self.fused.params.numel():
            raise ValueError(f"checkpoint holds {got['params'].numel()} parameters, the model {self.fused.params.numel()}")
        with torch.no_grad():
            self.fused.params.copy_(got["params"].to(self.device))
        self.fused.refresh_shadow()
        if got["m"] is not None:
            self.fused.m.copy_(got["m"].to(self.device))
            self.fused.v.copy_(got["v"].to(self.device))
            self.fused.set_adam_steps(got["step"])
        self.current_epoch = int(st["epoch"])
        self.current_angle = int(st["angle"])
        self.current_projection = int(st.get("projection", self.current_epoch * len(self.dataset) + self.current_angle))
        # the schedule at that projection (base_trainer.py:199-204): detector binning and points per ray
        dd, ppr = self.config["downsampling_detector"], self.config["points_per_ray"]
        P = self.current_projection
        self.downsample = max(1, int(dd["start"]) // 2 ** (P // int(dd["update_interval"])))
        if ppr.get("linear", True):
            self.points_per_ray = min(int(ppr["end"]), int(ppr["start"]) + P // int(ppr["update_interval"]))
        else:
            self.points_per_ray = min(int(ppr["end"]), int(ppr["start"]) * 2 ** (P // int(ppr["update_interval"])))
        # fit() draws one view order per epoch from this generator: replay the epochs already trained
        for _ in range(self.current_epoch):
            torch.randperm(len(self.dataset), generator=self._view_gen)
        if "lr_warmup" in st:
            self.lr_decay.load_state_dict(st["lr_scheduler"])
            self.lr_warmup.load_state_dict(st["lr_warmup"])
        else:
            # the reference's scheduler state has another layout: replay the per-projection steps
            for p in range(self.current_projection):
                self._lr_holder.step()
                self.lr_decay.step()
                if p + 1 <= self.config["warmup"]["steps"]:
                    self.lr_warmup.step()
        if got["lr"] is not None:
            self._lr_holder.param_groups[0]["lr"] = got["lr"]


class ProjectionsLoadedTrainer(BaseTrainer):
    """ProjectionsLoadedTrainer(config, projections, output_directory, checkpoint, ...) of
    nect/trainers/projections_loaded_trainer.py:17-31: projections given as an array [N, H, W], not rescaled."""

    def __init__(self, config: Config, projections, output_directory: str | Path | None = None,
                 checkpoint: str | Path | None = None, save_ckpt: bool = True, save_last: bool = True,
                 save_optimizer: bool = True, verbose: bool = True, log: bool = False, cancel_at: str | None = None,
                 keep_two: bool = True, prune: bool = True, *, data_range=None, dtype: torch.dtype = torch.float16,
                 device: str | torch.device = "cuda"):
        super().__init__(config, output_directory, checkpoint, save_ckpt, save_last, save_optimizer, verbose, log,
                         cancel_at, keep_two, prune, projections=projections, data_range=data_range, dtype=dtype,
                         device=device)


Trainer = ProjectionsLoadedTrainer
