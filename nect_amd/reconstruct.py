"""``reconstruct``: the public entry point of nect/reconstruct.py:113-250 on the fused HIP path.

Same arguments and quality presets (reconstruct.py:162-216).  Static mode
returns the volume rotated by 180 degrees in the (y, x) plane as the reference
does (reconstruct.py:247-248); dynamic mode returns the run directory.
Differences: ``dtype`` selects the compute precision (float16 by default) and
``model`` overrides the preset network (e.g. a smaller MLP); the default
dynamic model of the reference (QuadCubes composite) is not available yet,
dynamic mode uses a single 4-D hash grid.
"""
from __future__ import annotations

from pathlib import Path
from typing import Literal

import numpy as np
import torch

from .config import get_static_cfg, setup_cfg
from .sampling.geometry import Geometry
from .trainer import Trainer, load_projections


def reconstruct(geometry: Geometry, projections, mode: str = "static", angles=None, radians: bool = True,
                quality: Literal["poor", "low", "medium", "high", "higher", "highest"] = "high",
                niter: int | None = None, lr: float | None = None, timesteps=None, verbose: bool = True,
                log: bool = False, exp_name: str | None = None, flip_projections: bool = False,
                channel_order: str | None = None, config_override: dict | None = None, save_ckpt: bool = False,
                dtype: torch.dtype = torch.float16):
    cfg = get_static_cfg("hash_grid")
    if mode == "dynamic":
        cfg["mode"] = "dynamic"
    elif mode != "static":
        raise ValueError(f"unknown mode '{mode}'")
    # quality presets (reconstruct.py:173-204)
    if quality in ["poor", "low"]:
        cfg["loss"] = "L2"
        cfg["encoder"]["log2_hashmap_size"] = 19
        if quality == "poor":
            cfg["epochs"] = "0.02x"
            cfg["base_lr"] *= 10
        else:
            cfg["epochs"] = "0.1x"
            cfg["base_lr"] *= 5
    elif quality == "medium":
        cfg["epochs"] = "0.3x"
        cfg["base_lr"] *= 2
    elif quality == "high":
        cfg["epochs"] = "1x"
        cfg["base_lr"] /= 2
        cfg["warmup"]["steps"] = 500
    elif quality in ["higher", "highest"]:
        if quality == "higher":
            cfg["epochs"] = "2x"
            cfg["base_lr"] /= 2
            cfg["warmup"]["steps"] = 3000
            cfg["lr_scheduler"]["lrf"] = 0.05
        else:
            cfg["epochs"] = "4x"
            cfg["base_lr"] /= 5
            cfg["warmup"]["steps"] = 5000
            cfg["lr_scheduler"]["lrf"] = 0.01
            cfg["encoder"]["log2_hashmap_size"] = 23
        cfg["points_per_ray"]["end"] = "1.5x"
    if niter is not None:
        cfg["epochs"] = niter
    if lr is not None:
        cfg["base_lr"] = lr
    if angles is not None:
        geometry.set_angles(angles, radians)
    elif geometry.angles is None:
        raise ValueError("angles must be provided, either as an argument or in the `Geometry` object.")
    if timesteps is not None:
        geometry.set_timesteps(timesteps)
    if config_override is not None:
        for k, v in config_override.items():
            if isinstance(v, dict) and isinstance(cfg.get(k), dict):
                cfg[k].update(v)
            else:
                cfg[k] = v
    cfg["geometry"] = geometry
    config = setup_cfg(cfg)
    if isinstance(projections, (str, Path)):
        # files: channel order and flip as NeCTDataset applies them (nect/data.py:145-195)
        projections = load_projections(projections, geometry.nDetector, channel_order, flip_projections)
    elif flip_projections:
        projections = torch.flip(torch.as_tensor(projections), [-2, -1])
    out_dir = None
    if log or save_ckpt or mode == "dynamic":
        out_dir = Path("outputs") / (exp_name or "run")
    trainer = Trainer(config, projections, output_directory=out_dir, save_ckpt=save_ckpt,
                      save_last=(mode == "dynamic"), save_optimizer=False, verbose=verbose, log=log, dtype=dtype)
    trainer.fit()
    if mode == "static":
        vol = trainer.create_volume(save=False, cpu=True)
        return torch.rot90(vol, 2, (1, 2)).numpy()
    return out_dir
