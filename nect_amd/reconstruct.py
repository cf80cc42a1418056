"""``reconstruct`` and ``reconstruct_from_config_file``: the public entry points of nect/reconstruct.py:113-315
on the fused HIP path.

Same arguments and quality presets (reconstruct.py:162-216); ``config_override`` replaces top-level keys of
the configuration as the reference's ``cfg.update`` does (a nested dictionary replaces the whole entry).
Static mode returns the volume rotated by 180 degrees in the (y, x) plane as the reference does
(reconstruct.py:247-248); dynamic mode returns the run directory.  Differences: ``dtype`` selects the compute
precision (float16 by default).  Dynamic mode trains the reference's default QuadCubes model (four 3-D hash
grids over (z,y,x), (y,x,t), (x,z,t), (z,y,t) plus the coordinates, a 4 x 128 network).
"""
from __future__ import annotations

from pathlib import Path
from typing import Literal

import numpy as np
import torch

from .config import get_cfg, get_dynamic_cfg, get_static_cfg, setup_cfg
from .sampling.geometry import Geometry
from .trainer import BaseTrainer, ProjectionsLoadedTrainer, load_projections


def reconstruct(geometry: Geometry, projections, mode: str = "static", angles=None, radians: bool = True,
                quality: Literal["poor", "low", "medium", "high", "higher", "highest"] = "high",
                niter: int | None = None, lr: float | None = None, timesteps=None, verbose: bool = True,
                log: bool = False, exp_name: str | None = None, flip_projections: bool = False,
                channel_order: str | None = None, config_override: dict | None = None, save_ckpt: bool = False,
                dtype: torch.dtype = torch.float16):
    if mode == "static":
        cfg = get_static_cfg("hash_grid")
    elif mode == "dynamic":
        cfg = get_dynamic_cfg("quadcubes")            # reconstruct.py:164-165
    else:
        raise ValueError(f"unknown mode '{mode}'")
    if channel_order is not None:
        cfg["channel_order"] = channel_order
    cfg["flip"] = flip_projections
    from_files = isinstance(projections, (str, Path))
    cfg["img_path"] = str(projections) if from_files else "RECONSTRUCTING_FROM_ARRAY"
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
        cfg.update(config_override)                       # reconstruct.py:215-216
    cfg["geometry"] = geometry
    config = setup_cfg(cfg)
    out_dir = None
    if log or save_ckpt or mode == "dynamic":
        out_dir = Path("outputs") / (exp_name or "run")
    common = dict(output_directory=out_dir, save_ckpt=save_ckpt, save_last=(mode == "dynamic"),
                  save_optimizer=False, verbose=verbose, log=log, dtype=dtype)
    if from_files:
        # files: channel order, flip and range as NeCTDataset reads them (nect/data.py:99-197)
        trainer = BaseTrainer(config, **common)
    else:
        if flip_projections:
            projections = torch.flip(torch.as_tensor(projections), [-2, -1])
        trainer = ProjectionsLoadedTrainer(config, projections, **common)
    trainer.fit()
    if mode == "static":
        vol = trainer.create_volume(save=False, cpu=True)
        return torch.rot90(vol, 2, (1, 2)).numpy()
    return out_dir


def reconstruct_from_config_file(cfg: str | Path, log_path: str = "outputs", save_ckpt: bool = True,
                                 checkpoint: str | None = None, save_volume: bool = False, save_last: bool = True,
                                 save_optimizer: bool = True, cancel_at: str | None = None, prune: bool = True,
                                 keep_two: bool = True, dtype: torch.dtype = torch.float16):
    """Reconstruction from a config file (nect/reconstruct.py:253-315): the run goes to
    log_path/<the config file's folder>/<mode> (or the checkpoint's run directory when resuming); a
    continuous-scanning config trains with the accumulated step.  With save_volume the trained volume is
    written (create_volume) instead of training.  Returns the run directory."""
    config = get_cfg(cfg)
    folder = Path(cfg).resolve().parent.name
    exp = Path(log_path) / folder / config.mode
    if checkpoint:
        exp = Path(checkpoint).parent.parent
    trainer = BaseTrainer(config, output_directory=exp, checkpoint=checkpoint, save_ckpt=save_ckpt,
                          save_last=save_last, save_optimizer=save_optimizer, verbose=True, cancel_at=cancel_at,
                          keep_two=keep_two, prune=prune, dtype=dtype)
    if save_volume:
        exp.mkdir(parents=True, exist_ok=True)
        trainer.create_volume(save=True, save_path=str(exp / "final.npy"))
    else:
        trainer.fit()
    return exp
