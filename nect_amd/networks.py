"""NeCT's model classes (nect/network/networks.py) on this library's tiny-cuda-nn drop-in (nect_amd.tcnn).

  HashGrid    networks.py:28-65    one hash grid over (z, y, x); with include_identity a Composite of the
                                   grid and Identity over the same three coordinates
  QuadCubes   networks.py:392-449  the default dynamic model: four 3-D grids over (z,y,x), (y,x,t), (x,z,t),
                                   (z,y,t), plus Identity over (z,y,x,t) with include_identity
  HyperCubes  networks.py:451-508  the same four grids and a 4-D grid over (z,y,x,t), plus Identity
  HashGrid4D  a single 4-D grid over (z,y,x,t): the "n_dims_to_encode: 4" nested grid of HyperCubes alone

Each holds ``net``, a NetworkWithInputEncoding whose flat ``params`` is the model's only parameter (state-dict
key "net.params", as in the reference's checkpoints), and ``grid_sources``: for each nested hash grid, in order,
the sample coordinates (0 z, 1 y, 2 x, 3 t) its input dimensions read.  The nested grids of one model are encoded in one launch
per dimensionality (nect_grid_encode_parts); networks of 64 neurons run on the fully fused MFMA kernels, wider
ones (the dynamic presets' 128) as GEMMs.
"""
from __future__ import annotations

import torch

from .tcnn import NetworkWithInputEncoding


def _t_column(zyx: torch.Tensor, t) -> torch.Tensor:
    """torch.full((N, 1), t) as the reference concatenates the time (networks.py:498-499)."""
    return torch.full((zyx.size(0), 1), float(t), device=zyx.device, dtype=torch.float32)


class HashGrid(torch.nn.Module):
    def __init__(self, encoding_config: dict, network_config: dict, include_identity: bool = False,
                 dtype: torch.dtype = torch.float16, seed: int = 1337):
        super().__init__()
        self.include_identity = include_identity
        if include_identity:
            encoding = {"otype": "Composite", "nested": [{"n_dims_to_encode": 3, **encoding_config},
                                                         {"n_dims_to_encode": 3, "otype": "Identity"}]}
        else:
            encoding = encoding_config
        self.net = NetworkWithInputEncoding(3 + (3 if include_identity else 0), 1, encoding, network_config,
                                            seed=seed, dtype=dtype)
        self.grid_sources = [(0, 1, 2)]

    def forward(self, x: torch.Tensor, t=None) -> torch.Tensor:
        return self.net(torch.cat([x, x], dim=-1) if self.include_identity else x)


class HashGrid4D(torch.nn.Module):
    def __init__(self, encoding_config: dict, network_config: dict, dtype: torch.dtype = torch.float16,
                 seed: int = 1337):
        super().__init__()
        self.net = NetworkWithInputEncoding(4, 1, encoding_config, network_config, seed=seed, dtype=dtype)
        self.grid_sources = [(0, 1, 2, 3)]

    def forward(self, zyx: torch.Tensor, t) -> torch.Tensor:
        return self.net(torch.cat([zyx, _t_column(zyx, t)], dim=1))


class QuadCubes(torch.nn.Module):
    def __init__(self, encoding_config: dict, network_config: dict, include_identity: bool = True,
                 dtype: torch.dtype = torch.float16, seed: int = 1337):
        super().__init__()
        self.include_identity = include_identity
        nested = [{"n_dims_to_encode": 3, **encoding_config} for _ in range(4)]
        if include_identity:
            nested.append({"n_dims_to_encode": 4, "otype": "Identity"})
        self.net = NetworkWithInputEncoding(12 + (4 if include_identity else 0), 1,
                                            {"otype": "Composite", "nested": nested}, network_config, seed=seed,
                                            dtype=dtype)
        # (z,y,x), (y,x,t), (x,z,t), (z,y,t): the column order of forward()
        self.grid_sources = [(0, 1, 2), (1, 2, 3), (2, 0, 3), (0, 1, 3)]

    def forward(self, zyx: torch.Tensor, t) -> torch.Tensor:
        tc = _t_column(zyx, t)
        yxt = torch.cat([zyx[..., [1, 2]], tc], dim=1)
        xzt = torch.cat([zyx[..., [2, 0]], tc], dim=1)
        zyt = torch.cat([zyx[..., [0, 1]], tc], dim=1)
        cols = [zyx, yxt, xzt, zyt]
        if self.include_identity:
            cols.append(torch.cat([zyx, tc], dim=1))
        return self.net(torch.cat(cols, dim=-1))


class HyperCubes(torch.nn.Module):
    def __init__(self, encoding_config: dict, network_config: dict, include_identity: bool = True,
                 dtype: torch.dtype = torch.float16, seed: int = 1337):
        super().__init__()
        self.include_identity = include_identity
        nested = [{"n_dims_to_encode": 3, **encoding_config} for _ in range(4)]
        nested.append({"n_dims_to_encode": 4, **encoding_config})
        if include_identity:
            nested.append({"n_dims_to_encode": 4, "otype": "Identity"})
        self.net = NetworkWithInputEncoding(16 + (4 if include_identity else 0), 1,
                                            {"otype": "Composite", "nested": nested}, network_config, seed=seed,
                                            dtype=dtype)
        self.grid_sources = [(0, 1, 2), (1, 2, 3), (2, 0, 3), (0, 1, 3), (0, 1, 2, 3)]

    def forward(self, zyx: torch.Tensor, t) -> torch.Tensor:
        tc = _t_column(zyx, t)
        yxt = torch.cat([zyx[..., [1, 2]], tc], dim=1)
        xzt = torch.cat([zyx[..., [2, 0]], tc], dim=1)
        zyt = torch.cat([zyx[..., [0, 1]], tc], dim=1)
        zyxt = torch.cat([zyx, tc], dim=1)
        cols = [zyx, yxt, xzt, zyt, zyxt] + ([zyxt] if self.include_identity else [])
        return self.net(torch.cat(cols, dim=-1))


def build_model(config, dtype: torch.dtype = torch.float16) -> torch.nn.Module:
    """The model a configuration names (Config.get_model, nect/config.py:288-391): "hash_grid" (a 4-D grid in
    dynamic mode), "quadcubes", "hypercubes"."""
    name = str(config.raw.get("model") or "hash_grid").casefold()
    enc, net = config.encoder_config(), config.network_config()
    identity = bool(config.raw["net"].get("include_identity", False))
    if name == "hash_grid":
        if config.mode == "dynamic":
            return HashGrid4D(enc, net, dtype=dtype)
        return HashGrid(enc, net, include_identity=identity, dtype=dtype)
    if name == "quadcubes":
        return QuadCubes(enc, net, include_identity=identity, dtype=dtype)
    if name == "hypercubes":
        return HyperCubes(enc, net, include_identity=identity, dtype=dtype)
    raise NotImplementedError(f"model {name} is not implemented (hash_grid, quadcubes, hypercubes)")


__all__ = ["HashGrid", "HashGrid4D", "QuadCubes", "HyperCubes", "build_model"]
