"""Exports of a trained model on the fused query kernel.

Same entry points, arguments, outputs and file names as the reference:

  export_volume(base_path, binning, show_slices, ROIx, ROIy, ROIz)
      nect/static_export.py:15-144 -> base_path/volumefloat32/Static_Volume.tiff
  export_volumes(base_path, binning, avg_timesteps, timesteps_per_revolution, export_revolutions, show_slices,
                 ROIx, ROIy, ROIz, dtype)
      nect/dynamic_export.py:297-440 -> base_path/volumesfloat32/T{rev:04}_{proj:04}.tiff
  export_volumes_at_timesteps(base_path, timesteps, binning, ROIx, ROIy, ROIz, dtype)
      nect/dynamic_export.py:442-544 -> base_path/volumesfloat32/T{t:04}.tiff

A run directory holds config.yaml and checkpoints/last.ckpt as the trainer writes them (the reference's
layout).  The voxel grid is built exactly as the reference builds it -- the same ROI/binning arithmetic,
torch.linspace on the CPU for the static export and on the GPU for the dynamic ones -- and handed to
nect_query as explicit coordinate arrays; the network runs in the model's dtype (half by default, as
tiny-cuda-nn's output is half) and the value is rescaled in the kernel with the reference's three operations,
((mu / L_max) * (max - min)) + min, rounded as torch rounds them on the CPU (static) or the GPU (dynamic).
Where the reference averages several timesteps (export_volumes with avg_timesteps > 1) the accumulation runs
in torch with the reference's own operations.  Volumes are written as multi-page float32 TIFF files (one page
per z slice) with PIL.
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .config import Config, get_cfg
from .grid import GridDesc
from .mlp import MlpDesc
from .serving import RESCALE_DIVIDE, RESCALE_RECIPROCAL, QueryRequest, QueryServer
from .trainer import data_range_of_config, read_checkpoint_state


@dataclass
class ExportModel:
    """A trained hash-grid model on the device, ready for queries."""

    config: Config
    grid: GridDesc
    mlp: MlpDesc
    weights: torch.Tensor
    table: torch.Tensor
    maximum: float
    minimum: float

    @property
    def dtype(self) -> torch.dtype:
        return self.table.dtype

    @property
    def l_max(self) -> float:
        return float(self.config.geometry.max_distance_traveled)

    def server(self) -> QueryServer:
        return QueryServer(self.grid, self.mlp, self.weights, self.table, out_dtype=torch.float32, chunk=1 << 20,
                           max_jobs=64)

    def query(self, axes, t: float | None, rescale: int, mu_half: bool | None = None, identity: bool = False):
        """Values [nz, ny, nx] (float32, on the device) on the grid of the (z, y, x) coordinate tensors."""
        axes = [a.to(device=self.table.device, dtype=torch.float32).contiguous() for a in axes]
        n = tuple(int(a.numel()) for a in axes)
        half = self.dtype == torch.float16 if mu_half is None else mu_half
        kw = {} if identity else {"l_max": self.l_max, "span": self.maximum - self.minimum, "lo": self.minimum}
        r = QueryRequest(n, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.0 if t is None else float(t), mu_half=half,
                         rescale=rescale, axes=tuple(axes), **kw)
        return self.server().run_alone(r)


@dataclass
class ModuleExportModel:
    """A trained model the query kernel does not cover (composite encodings such as QuadCubes, networks wider
    than 64 neurons): evaluated as a module in chunks of 5,000,000 points (the reference's batch,
    static_export.py:113), its output in the compute dtype (half by default, as tiny-cuda-nn returns it), and
    rescaled with the reference's torch operations on the tensor the reference applies them to."""

    config: Config
    model: torch.nn.Module
    maximum: float
    minimum: float
    dtype: torch.dtype

    @property
    def l_max(self) -> float:
        return float(self.config.geometry.max_distance_traveled)

    def query(self, axes, t: float | None, rescale: int, mu_half: bool | None = None, identity: bool = False):
        dev = self.model.net.params.device
        axes = [a.to(device=dev, dtype=torch.float32) for a in axes]
        n = tuple(int(a.numel()) for a in axes)
        z, y, x = torch.meshgrid(axes, indexing="ij")
        grid = torch.stack((z.flatten(), y.flatten(), x.flatten()), 1)
        out = torch.empty(grid.shape[0], dtype=torch.float32, device=dev)
        with torch.no_grad():
            for s in range(0, grid.shape[0], 5_000_000):
                g = grid[s:s + 5_000_000].contiguous()
                mu = self.model(g, t) if self.config.mode == "dynamic" else self.model(g)
                out[s:s + g.shape[0]] = mu[:, 0].to(self.dtype).float()
        out = out.view(n)
        if identity:
            return out
        if rescale == RESCALE_DIVIDE:
            out = out.cpu()
        out = out / self.l_max
        return out * (self.maximum - self.minimum) + self.minimum


def load_model(base_path: str | Path, dtype: torch.dtype = torch.float16, device="cuda"):
    """The model of a run directory: config.yaml, checkpoints/last.ckpt (model parameters in the reference's
    checkpoint layout), and the projections' range (maximum, minimum) as NeCTDataset reports it.  A single
    hash grid with the 64-neuron network is served by the query kernel (ExportModel), any other model as a
    module (ModuleExportModel)."""
    from .mlp import MlpDesc
    from .networks import HashGrid, HashGrid4D, build_model

    base = Path(base_path)
    config = get_cfg(base / "config.yaml")
    st = torch.load(base / "checkpoints" / "last.ckpt", map_location="cpu", weights_only=True)
    params = read_checkpoint_state(st)["params"]
    maximum, minimum = data_range_of_config(config)
    module = build_model(config, dtype).to(device)
    if params.numel() != module.net.params.numel():
        raise ValueError(f"checkpoint holds {params.numel()} parameters, the configured model "
                         f"{module.net.params.numel()}")
    with torch.no_grad():
        module.net.params.copy_(params.to(device))
    net = module.net
    if not (isinstance(module, (HashGrid, HashGrid4D)) and len(net.parts) == 1 and isinstance(net.desc, MlpDesc)):
        return ModuleExportModel(config, module, maximum, minimum, dtype)
    enc, net = config.encoder_config(), config.network_config()
    n_dims = 4 if config.mode == "dynamic" else 3
    grid = GridDesc(n_dims, enc["n_levels"], enc["n_features_per_level"], enc["log2_hashmap_size"],
                    enc["base_resolution"], enc["per_level_scale"])
    mlp = MlpDesc(grid.n_output_dims, 1, net["n_neurons"], net["n_hidden_layers"], net["activation"],
                  net["output_activation"])
    p = params.to(device=device, dtype=dtype)
    return ExportModel(config, grid, mlp, p[: mlp.n_params], p[mlp.n_params:], maximum, minimum)


def export_extent(nVoxel, binning: int = 1, ROIx=None, ROIy=None, ROIz=None, sample_outside: int = 0):
    """((z_h, y_w, x_w), z range, y range, x range) of an export (nect/static_export.py:53-88, identical in
    nect/dynamic_export.py:485-516): the full grid is nVoxel[0] // binning by nVoxel[1] // binning twice;
    an ROI [i0, i1) of the grid padded by sample_outside voxels in y and x spans (i0 - rm) / n .. (i1 - rm) / n
    with (i1 - i0) // binning voxels (z unpadded)."""
    height, width = int(nVoxel[0]), int(nVoxel[1])
    z_h, y_w, x_w = height // binning, width // binning, width // binning
    rm = int(sample_outside)
    n_pad = [int(nVoxel[0]), int(nVoxel[1]) + 2 * rm, int(nVoxel[2]) + 2 * rm]
    xr, yr, zr = (0, 1), (0, 1), (0, 1)
    if ROIx is not None:
        xr = ((ROIx[0] - rm) / n_pad[2], (ROIx[1] - rm) / n_pad[2])
        x_w = (ROIx[1] - ROIx[0]) // binning
    if ROIy is not None:
        yr = ((ROIy[0] - rm) / n_pad[1], (ROIy[1] - rm) / n_pad[1])
        y_w = (ROIy[1] - ROIy[0]) // binning
    if ROIz is not None:
        zr = (ROIz[0] / n_pad[0], ROIz[1] / n_pad[0])
        z_h = (ROIz[1] - ROIz[0]) // binning
    return (z_h, y_w, x_w), zr, yr, xr


def _axes(extent, device) -> list[torch.Tensor]:
    (z_h, y_w, x_w), zr, yr, xr = extent
    return [torch.linspace(r[0], r[1], steps=n, dtype=torch.float32, device=device)
            for r, n in ((zr, z_h), (yr, y_w), (xr, x_w))]


def write_tiff(path: Path, volume: np.ndarray):
    """A [Z, Y, X] volume as a multi-page TIFF (one page per z slice)."""
    from PIL import Image

    vol = np.asarray(volume)
    if vol.dtype == np.float16 or vol.dtype == np.float64:
        vol = vol.astype(np.float32)
    pages = [Image.fromarray(np.ascontiguousarray(vol[k])) for k in range(vol.shape[0])]
    pages[0].save(path, save_all=True, append_images=pages[1:])


def read_tiff(path: Path) -> np.ndarray:
    """The pages of a multi-page TIFF stacked as [Z, Y, X]."""
    from PIL import Image, ImageSequence

    with Image.open(path) as im:
        return np.stack([np.array(p) for p in ImageSequence.Iterator(im)])


def _show_slices(m: ExportModel, base: Path, extent, t: float | None) -> Path:
    """Central z, y and x slices as PNG images (base/imgs/{z,y,x}.png), coordinates by torch.linspace on the
    GPU, the other axis at 0.5 (nect/static_export.py:89-105)."""
    from PIL import Image

    dev = "cuda"
    axes = _axes(extent, dev)
    half = torch.tensor([0.5], dtype=torch.float32, device=dev)
    (base / "imgs").mkdir(parents=True, exist_ok=True)
    for k, name in enumerate(["z", "y", "x"]):
        ax = list(axes)
        ax[k] = half
        out = m.query(ax, t, RESCALE_RECIPROCAL, identity=True).squeeze(k).cpu().numpy()
        lo, hi = float(out.min()), float(out.max())
        img = ((out - lo) / (hi - lo if hi > lo else 1.0) * 255).astype(np.uint8)
        Image.fromarray(img).save(base / "imgs" / f"{name}.png")
    return base / "imgs"


def export_volume(base_path: str | Path, binning: int = 1, show_slices: bool = False, ROIx=None, ROIy=None,
                  ROIz=None, dtype: torch.dtype = torch.float16) -> Path:
    """Static volume of a run (nect/static_export.py:15-144); returns the directory written to."""
    base = Path(base_path)
    m = load_model(base, dtype)
    g = m.config.geometry
    extent = export_extent(g.nVoxel, binning, ROIx, ROIy, ROIz, m.config.get("sample_outside", 0) or 0)
    if show_slices:
        return _show_slices(m, base, extent, None)
    out_dir = base / "volumefloat32"
    out_dir.mkdir(exist_ok=True, parents=True)
    # coordinates made on the CPU, as the reference makes them; values rescaled as on CPU tensors
    axes = _axes(extent, "cpu")
    vol = m.query(axes, None, RESCALE_DIVIDE).cpu().numpy()
    write_tiff(out_dir / "Static_Volume.tiff", vol)
    return out_dir


def projections_per_revolution(angles) -> int:
    """nect/dynamic_export.py:18-28: the first index where the angle increment changes by more than 0.01 rad
    (mod 2 pi) ends a revolution; all angles if it never does."""
    n = len(angles)
    for i in range(n - 2):
        d0 = (angles[i + 1] - angles[i]) % (2 * np.pi)
        d1 = (angles[i + 2] - angles[i + 1]) % (2 * np.pi)
        if abs(d1 - d0) > 0.01:
            return i + 2
    return n


def export_volumes_at_timesteps(base_path: str | Path, timesteps, binning: int = 1, ROIx=None, ROIy=None,
                                ROIz=None, dtype=np.float32, model_dtype: torch.dtype = torch.float16) -> Path:
    """Volumes of a dynamic run at the given times (nect/dynamic_export.py:442-544); returns the directory."""
    base = Path(base_path)
    m = load_model(base, model_dtype)
    if m.config.mode != "dynamic":
        raise ValueError("Only dynamic mode is supported for video creation")
    g = m.config.geometry
    extent = export_extent(g.nVoxel, binning, ROIx, ROIy, ROIz, m.config.get("sample_outside", 0) or 0)
    out_dir = base / "volumesfloat32"
    out_dir.mkdir(exist_ok=True, parents=True)
    ts = list(timesteps)
    if max(ts) > 1:
        ts = [t / len(g.angles) for t in ts]
    axes = _axes(extent, "cuda")
    for t in ts:
        vol = m.query(axes, float(t), RESCALE_RECIPROCAL).cpu().numpy().astype(dtype)
        write_tiff(out_dir / f"T{t:04}.tiff", vol)
    return out_dir


def export_volumes(base_path: str | Path, binning: int = 1, avg_timesteps: int = 1,
                   timesteps_per_revolution="all", export_revolutions="all", show_slices: bool = False, ROIx=None,
                   ROIy=None, ROIz=None, dtype=np.float32, model_dtype: torch.dtype = torch.float16) -> Path:
    """Volumes of a dynamic run per revolution (nect/dynamic_export.py:297-440); returns the directory."""
    base = Path(base_path)
    m = load_model(base, model_dtype)
    if m.config.mode != "dynamic":
        raise ValueError("Only dynamic mode is supported for video creation")
    g = m.config.geometry
    dev = torch.device("cuda", torch.cuda.current_device())
    extent = export_extent(g.nVoxel, binning, ROIx, ROIy, ROIz, m.config.get("sample_outside", 0) or 0)
    if show_slices:
        return _show_slices(m, base, extent, 0.5)
    out_dir = base / "volumesfloat32"
    out_dir.mkdir(exist_ok=True, parents=True)
    angles = list(g.angles)
    lin = torch.linspace(0, 1, steps=len(angles), device=dev)
    ppr = projections_per_revolution(angles)
    if export_revolutions == "all":
        export_revolutions = list(range(len(angles) // ppr))
    axes = _axes(extent, dev)
    for i in export_revolutions:
        r0, r1 = i * ppr, (i + 1) * ppr
        if timesteps_per_revolution != "all":
            sub = torch.linspace(lin[r0], lin[r1], timesteps_per_revolution)
        else:
            sub = lin[r0:r1]
        for j, t in enumerate(sub):
            if avg_timesteps == 1:
                # (mu / 1) rounded to half is mu: one kernel pass with the rescale on the GPU
                vol = m.query(axes, float(t), RESCALE_RECIPROCAL)
            else:
                acc = torch.zeros(tuple(int(a.numel()) for a in axes), device=dev)
                for avg in range(avg_timesteps):
                    tt = float((t + avg / len(angles)).float())
                    mu = m.query(axes, tt, RESCALE_RECIPROCAL, identity=True)
                    acc += mu.to(torch.float16 if m.dtype == torch.float16 else torch.float32) / avg_timesteps
                vol = acc / g.max_distance_traveled
                vol = vol * (m.maximum - m.minimum) + m.minimum
            out = vol.cpu().numpy().astype(dtype)
            proj_n = j if timesteps_per_revolution == "all" else (j * ppr) // timesteps_per_revolution
            write_tiff(out_dir / f"T{i:04}_{proj_n:04}.tiff", out)
    return out_dir
