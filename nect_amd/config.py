"""Training configuration: the subset of nect/config.py the hot path consumes.

Layering follows the reference (nect/config.py:579-689): defaults, then the
model preset, then user overrides, then the geometry.  Derived quantities keep
the reference's formulas:

  * per_level_scale = (max_resolution_factor * size / base_resolution)^(1/(L-1))
    (config.py:99-112), unless a ``per_level_scale`` is given directly;
  * lr = base_lr * sqrt(points_per_batch * devices / 1e6) unless ``lr`` is set
    (config.py:398-399);
  * loss selection by substring, "l1" tested first (config.py:481-497);
  * exponential/linear warm-up and cosine/exponential decay per projection
    (config.py:437-479), stepped with torch's own schedulers.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import torch
import yaml

from .sampling.geometry import Geometry

# nect/cfg/default.yaml (the keys the hot path reads)
DEFAULTS: dict[str, Any] = {
    "epochs": "auto",
    "loss": "L1",
    "optimizer": {"otype": "Adam", "weight_decay": 0.0, "beta1": 0.9, "beta2": 0.999},
    "base_lr": 0.001,
    "lr": None,
    "lr_scheduler": {"otype": "cosine", "lrf": "auto"},
    "warmup": {"steps": 500, "lr0": 0.001, "otype": "exponential"},
    "batch_per_proj": 1,
    "downsampling_detector": {"start": 2, "end": 1, "update_interval": 1000},
    "points_per_ray": {"start": 100, "end": "auto", "update_interval": "auto", "linear": True},
    "points_per_batch": "auto",
    "reconstruction_mode": "voxel",
    "uniform_ray_spacing": True,
    "clip_grad_value": None,
    "sample_outside": 0,
    "continous_scanning": False,
    "accumulation_steps": None,
}

# nect/cfg/static/hash_grid.yaml and the dynamic presets
PRESETS: dict[str, dict] = {
    "hash_grid": {
        "encoder": {"otype": "HashGrid", "n_levels": 19, "n_features_per_level": 4, "log2_hashmap_size": 21,
                    "base_resolution": 8, "max_resolution_factor": 2},
        "net": {"otype": "FullyFusedMLP", "activation": "LeakyReLU", "output_activation": "ReLU", "n_neurons": 64,
                "n_hidden_layers": 6},
        "model": "hash_grid",
        "mode": "static",
    },
}

# nect/cfg/dynamic/quadcubes.yaml and hypercubes.yaml
DYNAMIC_PRESETS: dict[str, dict] = {
    "quadcubes": {
        "encoder": {"otype": "HashGrid", "n_levels": 23, "n_features_per_level": 4, "log2_hashmap_size": 23,
                    "base_resolution": 16, "max_resolution_factor": 2},
        "net": {"otype": "FullyFusedMLP", "activation": "LeakyReLU", "output_activation": "None", "n_neurons": 128,
                "n_hidden_layers": 4, "include_identity": True},
        "base_lr": 0.0002,
        "batch_per_proj": 1,
        "warmup": {"steps": 2000, "lr0": 0.0001},
        "model": "quadcubes",
        "mode": "dynamic",
    },
    "hypercubes": {
        "encoder": {"otype": "HashGrid", "n_levels": 23, "n_features_per_level": 4, "log2_hashmap_size": 23,
                    "base_resolution": 16, "max_resolution_factor": 2},
        "net": {"otype": "FullyFusedMLP", "activation": "LeakyReLU", "output_activation": "None", "n_neurons": 128,
                "n_hidden_layers": 4, "include_identity": True},
        "base_lr": 0.0002,
        "batch_per_proj": 1,
        "warmup": {"steps": 2000, "lr0": 0.0001, "otype": "linear"},
        "model": "hypercubes",
        "mode": "dynamic",
    },
}


def _merge(base: dict, over: dict) -> dict:
    out = copy.deepcopy(base)
    for k, v in over.items():
        if isinstance(v, dict) and isinstance(out.get(k), dict):
            out[k] = _merge(out[k], v)
        else:
            out[k] = copy.deepcopy(v)
    return out


@dataclass
class Config:
    raw: dict
    geometry: Geometry
    devices: int = 1
    extras: dict = field(default_factory=dict)

    # ----------------------------------------------------------- accessors
    def __getitem__(self, k):
        return self.raw[k]

    def get(self, k, default=None):
        return self.raw.get(k, default)

    @property
    def mode(self) -> str:
        return self.raw.get("mode", "static")

    @property
    def encoder(self) -> dict:
        return self.raw["encoder"]

    @property
    def net(self) -> dict:
        return self.raw["net"]

    def per_level_scale(self) -> float:
        e = self.encoder
        if "per_level_scale" in e and e["per_level_scale"] is not None:
            return float(e["per_level_scale"])
        if e.get("max_resolution_type", "nDetector") == "nVoxel":
            nv = list(self.geometry.nVoxel)
            rm = self.raw.get("sample_outside", 0)
            size = max(nv[0], nv[1] + 2 * rm, nv[2] + 2 * rm)
        else:
            size = max(self.geometry.nDetector)
        return (e["max_resolution_factor"] * size / e["base_resolution"]) ** (1 / (e["n_levels"] - 1))

    def encoder_config(self) -> dict:
        e = self.encoder
        return {"otype": e.get("otype", "HashGrid"), "n_levels": e["n_levels"],
                "n_features_per_level": e["n_features_per_level"], "log2_hashmap_size": e["log2_hashmap_size"],
                "base_resolution": e["base_resolution"], "per_level_scale": self.per_level_scale()}

    def network_config(self) -> dict:
        n = self.net
        return {"otype": n.get("otype", "FullyFusedMLP"), "n_hidden_layers": n["n_hidden_layers"],
                "n_neurons": n["n_neurons"], "activation": n["activation"],
                "output_activation": n["output_activation"]}

    def points_per_batch(self) -> int:
        ppb = self.raw["points_per_batch"]
        if ppb == "auto":
            # the reference sizes from free memory and caps at 5M (config.py:380-390)
            ppb = 5_000_000
            self.raw["points_per_batch"] = ppb
        return int(ppb)

    def learning_rate(self) -> float:
        if self.raw.get("lr") is None:
            self.raw["lr"] = self.raw["base_lr"] * math.sqrt(self.points_per_batch() * self.devices / 1_000_000)
        return float(self.raw["lr"])

    def loss_name(self) -> str:
        key = str(self.raw["loss"]).casefold()
        if "l1" in key:
            return "L1"
        if "l2" in key:
            return "L2"
        raise NotImplementedError(f"Loss {self.raw['loss']} is not supported")

    def resolve_epochs(self, n_projections: int) -> int:
        ep = self.raw["epochs"]
        if isinstance(ep, str):
            base = math.ceil(49 / n_projections * max(self.geometry.nDetector))
            parts = ep.split("x")
            ep = math.ceil(base * float(parts[0])) if len(parts) == 2 and parts[0] else base
            self.raw["epochs"] = ep
        return int(ep)

    def schedulers(self, optim: torch.optim.Optimizer, n_angles: int):
        """(warmup, decay) torch LR schedulers as nect/config.py:437-479 builds them."""
        w = self.raw["warmup"]
        steps, lr0 = w["steps"], w["lr0"]
        if str(w.get("otype", "exponential")).casefold() == "linear":
            fac = (1 - lr0) / steps

            def warm(p):
                return p * fac if p < steps else 1
        else:
            fac = (1 / lr0) ** (1 / steps)

            def warm(p):
                return 1 / (fac ** (steps - p)) if p < steps else 1
        warmup = torch.optim.lr_scheduler.LambdaLR(optim, lr_lambda=warm)
        epochs = int(self.raw["epochs"])
        sch = self.raw["lr_scheduler"]
        if sch["lrf"] == "auto":
            sch["lrf"] = max(1e-2, min(0.5, 1 / (epochs * 0.1)))
        if "exponential" == str(sch["otype"]).casefold():
            gamma = sch["lrf"] ** (1 / (epochs * n_angles))
            decay = torch.optim.lr_scheduler.ExponentialLR(optim, gamma=gamma)
        else:
            decay = torch.optim.lr_scheduler.CosineAnnealingLR(optim, T_max=max(1, epochs * n_angles // self.devices),
                                                               eta_min=self.learning_rate() * sch["lrf"])
        return warmup, decay

    def to_yaml(self, path: str | Path):
        out = copy.deepcopy(self.raw)
        out["geometry"] = self.geometry.to_dict()
        with open(path, "w") as f:
            yaml.safe_dump(out, f)


def get_static_cfg(name: str = "hash_grid") -> dict:
    if name not in PRESETS:
        raise ValueError(f"unknown model preset '{name}'")
    return _merge(DEFAULTS, PRESETS[name])


def get_dynamic_cfg(name: str = "quadcubes") -> dict:
    """The defaults with a dynamic preset's entries replacing theirs (nect/config.py:596-610: dict.update)."""
    if name not in DYNAMIC_PRESETS:
        raise ValueError(f"unknown dynamic model preset '{name}'")
    cfg = copy.deepcopy(DEFAULTS)
    cfg.update(copy.deepcopy(DYNAMIC_PRESETS[name]))
    return cfg


def setup_cfg(cfg: dict) -> Config:
    cfg = copy.deepcopy(cfg)
    g = cfg.get("geometry")
    if isinstance(g, Geometry):
        geometry = g
    elif isinstance(g, dict):
        geometry = Geometry.from_dict(g, cfg.get("reconstruction_mode", "voxel"))
    else:
        raise ValueError("config needs a geometry")
    return Config(raw=cfg, geometry=geometry)


def get_cfg(path: str | Path) -> Config:
    """config.yaml (+ geometry inline or geometry.yaml beside it), layered over the defaults."""
    path = Path(path)
    with open(path) as f:
        user = yaml.safe_load(f)
    model = user.get("model", "hash_grid")
    if model in DYNAMIC_PRESETS and user.get("mode", "dynamic") == "dynamic":
        base = get_dynamic_cfg(model)
    elif model in PRESETS:
        base = get_static_cfg(model)
    else:
        base = DEFAULTS
    cfg = _merge(base, user)
    if cfg.get("geometry") is None or isinstance(cfg.get("geometry"), str):
        gpath = cfg.get("geometry") or (path.parent / "geometry.yaml")
        with open(gpath) as f:
            cfg["geometry"] = yaml.safe_load(f)
    return setup_cfg(cfg)
