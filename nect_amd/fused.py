"""Fused training step of the HIP path.

Replaces the body of BaseTrainer.fit's batch loop
(nect/trainers/base_trainer.py:517-653) for a hash-grid model:

    sampler      nect_ray_setup + nect_ray_tiles   (64-byte ray descriptors,
                                                    32-sample tiles; no [B,S,3]
                                                    point tensor, no mask/compact)
    encoder      nect_grid_encode_rays             features [slots, L*F] (dtype)
    MLP          nect_mlp_forward                  mu [slots] (float)
    integral     nect_integral_loss                y_pred [B], loss, dmu [slots]
    backward     nect_mlp_backward                 dW (float), dfeat [slots, L*F]
                 nect_ray_order +                  table gradient (float atomics over
                 nect_grid_backward_ray_groups     groups of neighbouring rays, coarse
                                                    levels summed per block in LDS)
    optimiser    nect_adam                         dense Adam + low-precision copy

Every launch is asynchronous on torch's current stream; the step never
synchronises the host.  In float16/bfloat16 the step follows the reference's
precision="16-mixed" (base_trainer.py:106-117): a dynamic loss scale held on
the device, a finite check of the gradient, and an optimiser step that is
skipped (scale halved) when the gradient is not finite -- decided on the
device (nect_grad_check + nect_adam_amp).  Parameters live in one float32 vector laid out as
[MLP weights | grid table] (the order tiny-cuda-nn's NetworkWithInputEncoding
keeps them in); for float16/bfloat16 a shadow copy in that dtype feeds the
kernels and is refreshed by the Adam kernel.
"""
from __future__ import annotations

import ctypes
import math

import torch

from . import _lib
from .grid import GridDesc
from .mlp import MlpDesc
from .sampling import ct_sampling as cs

LOSSES = {"l1": 0, "l2": 1}


# the line-integral kernel's mode that takes dL/dy per ray in place of a target (NECT_LOSS_GIVEN)
LOSS_GIVEN = 2


AGG_TABLE_BLOCKS = 512   # the scatter's LDS table (64-byte blocks)


def aggregation_plan(grid: GridDesc, geometry, step_mm: float, ray_fraction: float = 1.0) -> list[int]:
    """Per-level plan of nect_grid_backward_ray_groups: 0 (scatter directly) or the number R of LDS rounds.

    A ray group covers 8x8 neighbouring detector pixels and, per round, 16/R samples along the rays.  At level l
    (res_l cells across the volume) that region spans about (8 p c_xy + 1)(8 p c_z + 1)(d s c_xy + 1) cells
    (p the pixel footprint at the rotation axis, spread by 1/sqrt(ray_fraction) when the batch holds only that
    fraction of the pixels; s the sample spacing; d = 16/R samples; c the cells per mm), about a quarter of
    that many 64-byte blocks (eight x-neighbour corners share one; twice as many for a 4D grid, whose
    corners at one time t form two 3D sets, and for four features per entry, four entries to a block).  A
    level takes the fewest rounds (up to two) whose estimate fits 0.8 of the table, provided that leaves
    under 3 blocks per sample with two features (a direct scatter costs about 4.5 requests per sample) or
    0.8 with four (their LDS rounds cost twice the LDS atomics; measured on the QuadCubes preset: levels at
    1.0-1.5 blocks per sample scattered directly beat two LDS rounds, 78.0 against 81.7 ms per step).  Otherwise it is scattered directly: across rays (-1: an
    instruction holds one sample of 16 neighbouring rays, whose x-neighbour entries then share blocks; on the
    256^3 workload this beats one ray per instruction at every level it is used for) unless a cell is longer
    than two sample spacings, where one ray per instruction (0) sums runs of samples in one cell first.
    Every plan is correct; this one is fast."""
    dso = float(getattr(geometry, "distance_source_to_origin", 0.0) or 0.0)
    dod = float(getattr(geometry, "distance_origin_to_detector", 0.0) or 0.0)
    pitch = float(max(geometry.dDetector))
    foot = pitch * dso / (dso + dod) if dso > 0 and dod >= 0 else pitch
    foot /= math.sqrt(min(1.0, max(ray_fraction, 1e-6)))
    sv = getattr(geometry, "sVoxel", None)
    if sv is not None and min(sv) > 0:
        ext_z, ext_xy = float(sv[0]), math.sqrt(float(sv[1]) * float(sv[2]))
    else:
        r = float(getattr(geometry, "object_radius", 0.0) or 0.0)
        h = float(getattr(geometry, "object_height", 0.0) or 0.0)
        if r <= 0 or h <= 0:
            return [0] * grid.n_levels
        ext_z, ext_xy = h, 2 * r
    # a 4D grid's corners are two 3D sets (the batch's t falls in one cell of each level): twice the blocks;
    # a 64-byte block holds 16 / n_features entries: twice the blocks again with four features
    corner_sets = 2 ** max(0, grid.n_dims - 3) * max(1, grid.n_features // 2)
    max_blocks_per_sample = 3.0 if grid.n_features <= 2 else 0.8
    plan = []
    for res in grid.resolution:
        c_xy, c_z = res / ext_xy, res / ext_z
        cross = (8 * foot * c_xy + 1) * (8 * foot * c_z + 1)
        choice = 0
        for R in (1, 2):
            d = 16 // R
            est = cross * (d * step_mm * c_xy + 1) / 4 * corner_sets
            if est <= 0.8 * AGG_TABLE_BLOCKS:
                choice = R if est / (64 * d) < max_blocks_per_sample else 0
                break
        if choice == 0 and c_xy * step_mm >= 0.5:
            choice = -1
        plan.append(choice)
    return plan


def loss_code(name: str) -> int:
    # nect/config.py:481-497 tests "l1" first, so "L1+L2" selects L1
    key = name.casefold()
    if "l1" in key:
        return 0
    if "l2" in key:
        return 1
    raise NotImplementedError(f"Loss {name} is not supported")


class FusedTrainStep:
    """Owns the parameters, Adam state and per-step buffers of one hash-grid model."""

    def __init__(self, grid: GridDesc, mlp: MlpDesc, params: torch.Tensor, dtype: torch.dtype, max_rays: int,
                 max_points_per_ray: int, loss: str = "L2", betas=(0.9, 0.999), eps: float = 1e-8,
                 weight_decay: float = 0.0, memory_budget: int | None = None):
        """``memory_budget`` (bytes) bounds the per-sample buffers (features, their gradients, mu, dmu): a batch
        whose samples exceed it is processed in chunks of rays that fit, each chunk's gradient weighted by its
        share of the batch, and one optimiser step at the end."""
        if mlp.n_input != grid.n_output_dims:
            raise ValueError("MLP input width must equal the grid's L*F")
        if params.dtype != torch.float32 or not params.is_cuda:
            raise ValueError("params must be a float32 GPU tensor")
        self.grid = grid
        self.mlp = mlp
        self.dtype = dtype
        self.device = params.device
        self.n_mlp = mlp.n_params
        self.params = params
        if params.numel() != self.n_mlp + grid.n_params:
            raise ValueError("params must hold the MLP weights followed by the grid table")
        self.grad = torch.zeros_like(params)
        self.m = torch.zeros_like(params)
        self.v = torch.zeros_like(params)
        self.shadow = params if dtype == torch.float32 else params.to(dtype)
        self.loss = loss_code(loss)
        self.betas = betas
        self.eps = eps
        self.weight_decay = weight_decay
        # clip_grad_value of the config (base_trainer.py:651-652): gradients clamped to +-value before Adam
        self.clip_grad_value: float | None = None
        self.step_count = 0
        # mixed precision (16-bit dtypes): the NectAmpState on the device; its scale is initialised at the first
        # step from the batch shape (see _amp_init), grows by amp_growth_factor after amp_growth_interval finite
        # steps and backs off by amp_backoff_factor after a non-finite one (torch.amp.GradScaler's defaults)
        self.amp = dtype != torch.float32
        self.amp_state = torch.zeros(_lib.AMP_STATE_BYTES // 4, dtype=torch.int32, device=self.device)
        self.amp_ready = False
        self.amp_growth_factor = 2.0
        self.amp_backoff_factor = 0.5
        self.amp_growth_interval = 2000
        self._amp_pending_step: int | None = None
        self.profile = False
        self._timers: dict[str, list] = {}
        # features between encoder and MLP in the group-planar layout when the grid allows it (each XCD then
        # writes whole rows of its own plane); the MLP reads that layout directly
        self.grouped = grid.n_levels % 8 == 0 and mlp.n_input == mlp.n_input_padded
        self.memory_budget = memory_budget
        # table-gradient schedule: "groups" (nect_ray_order + nect_grid_backward_ray_groups) or "tiles"
        # (nect_grid_backward_rays in batch order); n_agg_levels None picks aggregated_levels()
        self.scatter = "groups"
        # the grouped schedule's patch order applied to the whole step (see _presort)
        self.presort = True
        self._y_sorted = torch.empty(0, dtype=torch.float32, device=self.device)
        self.agg_rounds: list[int] | None = None   # per level; None: aggregation_plan()
        self.scatter_split = False
        self._order_ws = torch.empty(0, dtype=torch.uint8, device=self.device)
        self._order_need: dict = {}
        self._plan_cache: dict = {}
        self._chunk_counts: list = []
        self._alloc(max_rays, max_points_per_ray)

    def _x_groups(self):
        return self.grid.struct_ref() if self.grouped else None

    # ------------------------------------------------------------ timing

    def _call(self, label: str, entry: str, *args):
        """Launches one entry point; with ``profile`` set, brackets it with events on the launch stream."""
        if not self.profile:
            _lib.call(entry, *args)
            return
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        _lib.call(entry, *args)
        e1.record()
        self._timers.setdefault(label, []).append((e0, e1))

    def reset_timers(self):
        self._timers = {}

    def timer_means(self) -> dict[str, float]:
        """Mean duration (ms) of each timed launch (synchronises)."""
        torch.cuda.synchronize(self.device)
        return {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in self._timers.items() if v}

    def last_samples(self) -> int:
        """Kept samples of the last batch (synchronises)."""
        if self._chunk_counts:
            return int(sum(int(c.item()) for c in self._chunk_counts))
        _, _, count, _ = cs.ray_fields(self.rays[: self._last_rays])
        return int(count.sum().item())

    def slot_bytes(self) -> int:
        """Bytes of the per-sample buffers per slot."""
        F = self.grid.n_output_dims
        esz = torch.empty((), dtype=self.dtype).element_size()
        return F * esz + F * 4 + 8

    def _alloc(self, max_rays: int, S: int):
        dev = self.device
        self.max_rays = max_rays
        self.max_S = S
        per_ray = self.slot_bytes() * cs.tiles_needed(1, S) * 32
        chunk = max_rays if self.memory_budget is None else max(1, min(max_rays, self.memory_budget // per_ray))
        self.chunk_rays = chunk
        self.max_tiles = cs.tiles_needed(chunk, S)
        slots = self.max_tiles * 32
        F = self.grid.n_output_dims
        self.rays = torch.empty((chunk, _lib.RAY_BYTES), dtype=torch.uint8, device=dev)
        self.order = torch.empty(chunk, dtype=torch.int32, device=dev)
        self.tile_offsets = torch.empty(chunk + 1, dtype=torch.int32, device=dev)
        self.tile_ray = torch.empty(self.max_tiles, dtype=torch.int32, device=dev)
        self.n_tiles = torch.empty(1, dtype=torch.int32, device=dev)
        self.features = torch.empty((slots, F), dtype=self.dtype, device=dev)
        self.mu = torch.empty(slots, dtype=torch.float32, device=dev)
        self.dmu = torch.empty(slots, dtype=torch.float32, device=dev)
        self.dfeat = torch.empty((slots, F), dtype=torch.float32, device=dev)
        self.y_pred = torch.empty(max_rays, dtype=torch.float32, device=dev)
        self.partials = torch.empty((max_rays + 3) // 4, dtype=torch.float32, device=dev)
        self.loss_value = torch.empty(1, dtype=torch.float32, device=dev)
        ws = int(_lib.lib().nect_mlp_backward_workspace(self.mlp.struct_ref(), _lib.dtype_code(self.dtype), slots))
        self.workspace = torch.empty(max(ws, 16), dtype=torch.uint8, device=dev)

    @property
    def mlp_weights(self) -> torch.Tensor:
        return self.shadow[: self.n_mlp]

    @property
    def table(self) -> torch.Tensor:
        return self.shadow[self.n_mlp:]

    def forward_rays(self, geometry, angle: float, ray_index: torch.Tensor, starting_ray_index: int, num_rays: int,
                     points_per_ray: int, step_mm: float, uniform: bool, jitter: torch.Tensor,
                     detector_offset: torch.Tensor, t: float = 0.0):
        """Sampler + encoder + MLP for one batch (at most ``chunk_rays`` rays); leaves mu [slots] in slot order."""
        if num_rays > self.max_rays or points_per_ray > self.max_S:
            self._alloc(max(num_rays, self.max_rays), max(points_per_ray, self.max_S))
        if num_rays > self.chunk_rays:
            raise ValueError(f"forward_rays takes at most {self.chunk_rays} rays under the memory budget")
        st = _lib.stream_of(self.device)
        rays = self.rays[:num_rays]
        self._last_rays = num_rays
        g = geometry._struct()
        self._call("ray_setup", "nect_ray_setup", g, float(angle), _lib.ptr(ray_index), ray_index.numel(), int(num_rays),
                  int(starting_ray_index), int(points_per_ray), float(step_mm), int(bool(uniform)), _lib.ptr(jitter),
                  _lib.ptr(detector_offset), _lib.ptr(rays), st)
        max_tiles = cs.tiles_needed(num_rays, points_per_ray)
        self._call("ray_tiles", "nect_ray_tiles", _lib.ptr(rays), int(num_rays), int(max_tiles), _lib.ptr(self.tile_offsets),
                  _lib.ptr(self.tile_ray), _lib.ptr(self.n_tiles), st)
        F = self.grid.n_output_dims
        self._call("grid_encode", "nect_grid_encode_rays", self.grid.struct_ref(), _lib.dtype_code(self.dtype), _lib.ptr(self.table),
                  _lib.ptr(rays), _lib.ptr(self.tile_offsets), _lib.ptr(self.tile_ray), _lib.ptr(self.n_tiles),
                  int(max_tiles), float(t), _lib.ptr(self.features), F, int(self.grouped), st)
        slots = max_tiles * 32
        self._call("mlp_forward", "nect_mlp_forward", self.mlp.struct_ref(), _lib.dtype_code(self.dtype), _lib.ptr(self.mlp_weights),
                  _lib.ptr(self.features), slots, F, self._x_groups(), _lib.ptr(self.n_tiles), _lib.ptr(self.mu), 1, st)
        return max_tiles

    def step(self, geometry, angle: float, ray_index: torch.Tensor, starting_ray_index: int, num_rays: int,
             points_per_ray: int, step_mm: float, uniform: bool, jitter: torch.Tensor, detector_offset: torch.Tensor,
             target: torch.Tensor, l_max: float, lr: float, t: float = 0.0, optimizer_step: bool = True) -> torch.Tensor:
        """One full training step; returns the (device) mean loss of the batch."""
        if num_rays > self.max_rays or points_per_ray > self.max_S:
            self._alloc(max(num_rays, self.max_rays), max(points_per_ray, self.max_S))
        self._chunk_counts = []
        if num_rays <= self.chunk_rays:
            self._chunk(geometry, angle, ray_index, starting_ray_index, num_rays, points_per_ray, step_mm, uniform,
                        jitter, detector_offset, target, self.y_pred, l_max, t, 1.0)
            total = self.loss_value
        else:
            total = torch.zeros(1, dtype=torch.float32, device=self.device)
            for c0 in range(0, num_rays, self.chunk_rays):
                n = min(self.chunk_rays, num_rays - c0)
                share = n / num_rays
                self._chunk(geometry, angle, ray_index, starting_ray_index + c0, n, points_per_ray, step_mm, uniform,
                            jitter[c0:c0 + n], detector_offset[c0:c0 + n], target[c0:c0 + n], self.y_pred[c0:], l_max, t,
                            share)
                _, _, count, _ = cs.ray_fields(self.rays[:n])
                self._chunk_counts.append(count.sum())
                total += self.loss_value * share
            self.loss_value.copy_(total)
        if optimizer_step:
            self.adam(lr)
        return self.loss_value

    def _chunk(self, geometry, angle, ray_index, start, num_rays, points_per_ray, step_mm, uniform, jitter,
               detector_offset, target, y_pred, l_max, t, share):
        """Forward, loss and backward of num_rays rays; gradients are added with weight `share`."""
        perm = self._presort(geometry, ray_index, start, num_rays, points_per_ray)
        y_out = y_pred
        if perm is not None:
            ray_index, start = self._sorted_index(geometry, ray_index, start, num_rays, perm), 0
            jitter, detector_offset = jitter[:num_rays][perm], detector_offset[:num_rays][perm]
            target = target[:num_rays][perm].contiguous()
            y_pred = self._y_sorted
        max_tiles = self.forward_rays(geometry, angle, ray_index, start, num_rays, points_per_ray, step_mm, uniform,
                                      jitter, detector_offset, t)
        st = _lib.stream_of(self.device)
        rays = self.rays[:num_rays]
        slots = max_tiles * 32
        F = self.grid.n_output_dims
        self._call("integral_loss", "nect_integral_loss", _lib.ptr(rays), _lib.ptr(self.tile_offsets), int(num_rays),
                   _lib.ptr(self.mu), _lib.ptr(target), float(l_max), self.loss, _lib.ptr(y_pred), _lib.ptr(self.dmu),
                   _lib.ptr(self.partials), _lib.ptr(self.loss_value), st)
        if perm is not None:
            y_out[:num_rays].index_copy_(0, perm, y_pred[:num_rays])
        loss_scale, loss_scale_dev = self._loss_scale(num_rays, points_per_ray)
        self._call("mlp_backward", "nect_mlp_backward", self.mlp.struct_ref(), _lib.dtype_code(self.dtype),
                   _lib.ptr(self.mlp_weights), _lib.ptr(self.features), slots, F, self._x_groups(),
                   _lib.ptr(self.n_tiles), _lib.ptr(self.dmu), 1, loss_scale, loss_scale_dev, float(share),
                   _lib.ptr(self.grad), _lib.ptr(self.dfeat), F, _lib.ptr(self.workspace), self.workspace.numel(), st)
        if perm is None:
            self._ray_order(geometry, ray_index, start, num_rays, points_per_ray)
        self._grid_backward(geometry, num_rays, points_per_ray, step_mm, max_tiles, t, ordered=perm is None)

    def _presort(self, geometry, ray_index, start: int, num_rays: int, points_per_ray: int):
        """With ``presort`` and the grouped schedule: the patch order of the batch (int64), which the whole step
        then follows -- the encoder's gathers of neighbouring rays share L2 lines, the scatter's groups are
        consecutive rays; per-ray outputs are returned in the caller's order.  None otherwise."""
        if not (self.presort and self._use_groups(num_rays, points_per_ray)):
            return None
        self._ray_order(geometry, ray_index, start, num_rays, points_per_ray)
        if self._y_sorted.numel() < num_rays:
            self._y_sorted = torch.empty(num_rays, dtype=torch.float32, device=self.device)
        return self.order[:num_rays].long()

    def _sorted_index(self, geometry, ray_index, start: int, num_rays: int, perm):
        """ray_index[start + perm] for the batch (slots past the index get a pixel outside the detector, which
        the sampler leaves empty, as it does for the slots themselves)."""
        n = ray_index.numel()
        have = max(0, min(num_rays, n - start))
        if have == num_rays:
            src = ray_index[start:start + num_rays]
        else:
            src = torch.full((num_rays,), geometry.nDetector[0] * geometry.nDetector[1], dtype=ray_index.dtype,
                             device=ray_index.device)
            if have > 0:
                src[:have] = ray_index[start:start + have]
        return src[perm].contiguous()

    # the grouped scatter takes one workgroup per 64 rays x 16 samples; below this many workgroups it would
    # leave the device idle, and the tile schedule (one wave per 32 samples) is faster
    GROUPS_MIN_WORKGROUPS = 2048

    def _use_groups(self, num_rays: int, points_per_ray: int) -> bool:
        passes = 2 * ((points_per_ray + 31) // 32)
        return self.scatter == "groups" and ((num_rays + 63) // 64) * passes >= self.GROUPS_MIN_WORKGROUPS

    def _ray_order(self, geometry, ray_index: torch.Tensor, start: int, num_rays: int, points_per_ray: int):
        """Groups the batch's rays by detector patch for the grouped scatter (self.order)."""
        if not self._use_groups(num_rays, points_per_ray):
            return
        g = geometry._struct()
        key = tuple(geometry.nDetector)
        need = self._order_need.get(key)
        if need is None:
            need = int(_lib.lib().nect_ray_order_workspace(g))
            if need < 0:
                raise RuntimeError(f"nect_ray_order_workspace failed: {_lib.lib().nect_last_error().decode()}")
            self._order_need[key] = need
        if self._order_ws.numel() < need:
            self._order_ws = torch.empty(need, dtype=torch.uint8, device=self.device)
        self._call("ray_order", "nect_ray_order", g, _lib.ptr(ray_index), ray_index.numel(), int(num_rays), int(start),
                   _lib.ptr(self.order), _lib.ptr(self._order_ws), self._order_ws.numel(), _lib.stream_of(self.device))

    def _plan(self, geometry, step_mm: float, num_rays: int):
        """The scatter's per-level plan as a ctypes array (cached per geometry, spacing and batch size)."""
        if self.agg_rounds is not None:
            return (ctypes.c_int32 * self.grid.n_levels)(*self.agg_rounds)
        key = (tuple(geometry.nDetector), tuple(geometry.dDetector),
               getattr(geometry, "distance_source_to_origin", None), getattr(geometry, "distance_origin_to_detector", None),
               tuple(getattr(geometry, "sVoxel", ()) or ()), getattr(geometry, "object_radius", None),
               getattr(geometry, "object_height", None), float(step_mm), int(num_rays))
        plan = self._plan_cache.get(key)
        if plan is None:
            n_pix = geometry.nDetector[0] * geometry.nDetector[1]
            levels = aggregation_plan(self.grid, geometry, step_mm, num_rays / max(1, n_pix))
            plan = (ctypes.c_int32 * self.grid.n_levels)(*levels)
            if len(self._plan_cache) > 64:
                self._plan_cache.clear()
            self._plan_cache[key] = plan
        return plan

    def _grid_backward(self, geometry, num_rays: int, points_per_ray: int, step_mm: float, max_tiles: int,
                       t: float, ordered: bool = True, plan=None):
        """Table gradient of the sampled batch from dfeat (rays, tiles and, when ``ordered``, self.order set;
        else the rays are already in patch order); ``plan`` overrides the batch's own aggregation plan."""
        st = _lib.stream_of(self.device)
        F = self.grid.n_output_dims
        rays = self.rays[:num_rays]
        grad = self.grad[self.n_mlp:].data_ptr()
        if self._use_groups(num_rays, points_per_ray):
            self._call("grid_backward", "nect_grid_backward_ray_groups", self.grid.struct_ref(), _lib.ptr(rays),
                       _lib.ptr(self.tile_offsets), int(num_rays), _lib.ptr(self.order) if ordered else None,
                       int(points_per_ray), float(t), _lib.ptr(self.dfeat), F, grad,
                       plan if plan is not None else self._plan(geometry, step_mm, num_rays),
                       int(bool(self.scatter_split)), st)
        else:
            self._call("grid_backward", "nect_grid_backward_rays", self.grid.struct_ref(), _lib.ptr(rays),
                       _lib.ptr(self.tile_offsets), _lib.ptr(self.tile_ray), _lib.ptr(self.n_tiles),
                       int(max_tiles), float(t), _lib.ptr(self.dfeat), F, grad, st)

    def step_accumulated(self, geometry, angles, ray_index: torch.Tensor, starting_ray_index: int, num_rays: int,
                         points_per_ray: int, step_mm: float, uniform: bool, jitters, detector_offsets,
                         target: torch.Tensor, l_max: float, lr: float, t: float = 0.0,
                         optimizer_step: bool = True) -> torch.Tensor:
        """Continuous-scanning step (nect/trainers/continous_scanning_trainer.py:81-167): each ray's prediction
        is the mean of its line integrals at the K sub-angles ``angles`` (one jitter and detector offset per
        sub-angle), accumulated as sum_k (y_k / K) in sub-angle order; the loss compares that mean with the
        target.  Pass 1 evaluates the K line integrals; pass 2 recomputes each sub-angle's samples and
        features (bit-identical) and back-propagates dL/dy / K through it, so memory does not grow with K."""
        K = len(angles)
        if K < 1 or len(jitters) != K or len(detector_offsets) != K:
            raise ValueError("one jitter and one detector offset per sub-angle")
        if num_rays > self.max_rays or points_per_ray > self.max_S:
            self._alloc(max(num_rays, self.max_rays), max(points_per_ray, self.max_S))
        if num_rays > self.chunk_rays:
            raise ValueError("the accumulated step runs in one pass; raise the memory budget")
        self._chunk_counts = []
        st = _lib.stream_of(self.device)
        perm = self._presort(geometry, ray_index, starting_ray_index, num_rays, points_per_ray)
        if perm is not None:
            ray_index, starting_ray_index = self._sorted_index(geometry, ray_index, starting_ray_index, num_rays,
                                                               perm), 0
            jitters = [j[:num_rays][perm] for j in jitters]
            detector_offsets = [o[:num_rays][perm] for o in detector_offsets]
            target = target[:num_rays][perm]
        ys = torch.empty((K, num_rays), dtype=torch.float32, device=self.device)
        zeros = torch.zeros(num_rays, dtype=torch.float32, device=self.device)
        rays = self.rays[:num_rays]
        for k in range(K):
            self.forward_rays(geometry, angles[k], ray_index, starting_ray_index, num_rays, points_per_ray, step_mm,
                              uniform, jitters[k], detector_offsets[k], t)
            self._call("integral_loss", "nect_integral_loss", _lib.ptr(rays), _lib.ptr(self.tile_offsets),
                       int(num_rays), _lib.ptr(self.mu), _lib.ptr(zeros), float(l_max), LOSS_GIVEN, _lib.ptr(ys[k]),
                       _lib.ptr(self.dmu), _lib.ptr(self.partials), _lib.ptr(self.loss_value), st)
        y = ys[0] / K
        for k in range(1, K):
            y = y + ys[k] / K
        if perm is None:
            self.y_pred[:num_rays].copy_(y)
        else:
            self.y_pred[:num_rays].index_copy_(0, perm, y)
        diff = y - target[:num_rays]
        if self.loss == 1:
            loss = torch.mean(diff * diff)
            gy = 2.0 * diff / num_rays
        else:
            loss = torch.mean(diff.abs())
            gy = torch.sign(diff) / num_rays
        gy = gy.contiguous()
        F = self.grid.n_output_dims
        if perm is None:
            self._ray_order(geometry, ray_index, starting_ray_index, num_rays, points_per_ray)
        for k in range(K):
            max_tiles = self.forward_rays(geometry, angles[k], ray_index, starting_ray_index, num_rays,
                                          points_per_ray, step_mm, uniform, jitters[k], detector_offsets[k], t)
            slots = max_tiles * 32
            self._call("integral_loss", "nect_integral_loss", _lib.ptr(rays), _lib.ptr(self.tile_offsets),
                       int(num_rays), _lib.ptr(self.mu), _lib.ptr(gy), float(l_max), LOSS_GIVEN, _lib.ptr(ys[k]),
                       _lib.ptr(self.dmu), _lib.ptr(self.partials), _lib.ptr(self.loss_value), st)
            loss_scale, loss_scale_dev = self._loss_scale(num_rays, points_per_ray)
            self._call("mlp_backward", "nect_mlp_backward", self.mlp.struct_ref(), _lib.dtype_code(self.dtype),
                       _lib.ptr(self.mlp_weights), _lib.ptr(self.features), slots, F, self._x_groups(),
                       _lib.ptr(self.n_tiles), _lib.ptr(self.dmu), 1, loss_scale, loss_scale_dev, 1.0 / K,
                       _lib.ptr(self.grad), _lib.ptr(self.dfeat), F, _lib.ptr(self.workspace),
                       self.workspace.numel(), st)
            self._grid_backward(geometry, num_rays, points_per_ray, step_mm, max_tiles, t, ordered=perm is None)
        self.loss_value.copy_(loss.reshape(1))
        if optimizer_step:
            self.adam(lr)
        return self.loss_value

    def predict(self, num_rays: int, l_max: float) -> torch.Tensor:
        """Line integrals y_pred [num_rays] of the batch that forward_rays sampled last (loss terms unused)."""
        zeros = torch.zeros(num_rays, dtype=torch.float32, device=self.device)
        _lib.call("nect_integral_loss", _lib.ptr(self.rays[:num_rays]), _lib.ptr(self.tile_offsets), int(num_rays),
                  _lib.ptr(self.mu), _lib.ptr(zeros), float(l_max), self.loss, _lib.ptr(self.y_pred),
                  _lib.ptr(self.dmu), _lib.ptr(self.partials), _lib.ptr(self.loss_value), _lib.stream_of(self.device))
        return self.y_pred[:num_rays].clone()

    # ------------------------------------------------------------ precision

    @staticmethod
    def static_loss_scale(num_rays: int, points_per_ray: int) -> float:
        """dL/dmu ~ (y_pred - target) * distance / (l_max * B): B * S / 2 undoes the mean over the batch and the
        1/S sample spacing, so scaled values are O(|y_pred - target|)."""
        return 0.5 * num_rays * points_per_ray

    def _amp_init(self, num_rays: int, points_per_ray: int):
        """Initial loss scale of the 16-bit step: the static scale of the first batch rounded to a power of two
        (then dividing by it is exact)."""
        s = self.static_loss_scale(num_rays, points_per_ray)
        self.init_amp(2.0 ** round(math.log2(max(s, 1.0))))

    def _loss_scale(self, num_rays: int, points_per_ray: int):
        """(loss_scale, loss_scale_dev) arguments of nect_mlp_backward."""
        if not self.amp:
            return float(self.static_loss_scale(num_rays, points_per_ray)), None
        if not self.amp_ready:
            self._amp_init(num_rays, points_per_ray)
        return 1.0, _lib.ptr(self.amp_state)

    def amp_status(self) -> dict:
        """The device-side mixed-precision state (synchronises): scale, Adam steps taken, steps skipped."""
        st = self.amp_state.cpu()
        return {"scale": float(st[:1].view(torch.float32).item()), "growth_tracker": int(st[1]),
                "step": int(st[2]), "found_inf": int(st[3]), "skipped": int(st[4])}

    def adam_steps(self) -> int:
        """Optimiser steps taken (16-bit: without the skipped ones; synchronises)."""
        if self.amp and self.amp_ready:
            return self.amp_status()["step"]
        if self.amp and self._amp_pending_step is not None:
            return self._amp_pending_step
        return self.step_count

    def set_adam_steps(self, n: int):
        """Sets the optimiser step count (resuming from a checkpoint; applied when the loss scale is set up)."""
        self.step_count = int(n)
        if self.amp:
            if self.amp_ready:
                self.amp_state[2] = int(n)
            else:
                self._amp_pending_step = int(n)

    def init_amp(self, scale: float):
        """Sets the 16-bit step's loss scale explicitly (instead of from the first batch)."""
        if self.amp:
            _lib.call("nect_amp_init", _lib.ptr(self.amp_state), float(scale), _lib.stream_of(self.device))
            self.amp_ready = True
            if self._amp_pending_step is not None:
                self.amp_state[2] = self._amp_pending_step
                self._amp_pending_step = None

    def adam(self, lr: float):
        self.step_count += 1
        st = _lib.stream_of(self.device)
        if self.amp:
            if not self.amp_ready:
                raise RuntimeError("no backward pass ran before the optimiser step")
            self._call("grad_check", "nect_grad_check", _lib.ptr(self.grad), self.grad.numel(), _lib.ptr(self.amp_state),
                       st)
            self._call("adam", "nect_adam_amp", _lib.ptr(self.params), _lib.ptr(self.grad), _lib.ptr(self.m),
                       _lib.ptr(self.v), _lib.ptr(self.shadow), _lib.dtype_code(self.dtype), self.params.numel(),
                       float(lr), float(self.betas[0]), float(self.betas[1]), float(self.eps), float(self.weight_decay),
                       _lib.ptr(self.amp_state), float(self.amp_growth_factor), float(self.amp_backoff_factor),
                       int(self.amp_growth_interval), float(self.clip_grad_value or 0.0), st)
            return
        self._call("adam", "nect_adam", _lib.ptr(self.params), _lib.ptr(self.grad), _lib.ptr(self.m), _lib.ptr(self.v),
                  None, _lib.dtype_code(self.dtype), self.params.numel(), float(lr), float(self.betas[0]),
                  float(self.betas[1]), float(self.eps), float(self.weight_decay), self.step_count,
                  float(self.clip_grad_value or 0.0), st)

    def refresh_shadow(self):
        """Re-derive the low-precision copy after the float32 parameters were changed outside Adam."""
        if self.dtype != torch.float32:
            _lib.call("nect_cast", _lib.ptr(self.params), _lib.ptr(self.shadow), _lib.dtype_code(self.dtype),
                      self.params.numel(), _lib.stream_of(self.device))


def xavier_init(mlp: MlpDesc, generator: torch.Generator) -> torch.Tensor:
    """MLP weights, uniform(+-sqrt(6/(fan_in+fan_out))) per layer; padded rows/cols zero."""
    parts = []
    for i, (r, c) in enumerate(mlp.layer_shapes()):
        fan_in = mlp.n_input if i == 0 else c
        fan_out = mlp.n_output if i == len(mlp.layer_shapes()) - 1 else r
        bound = math.sqrt(6.0 / (fan_in + fan_out))
        w = (torch.rand((r, c), generator=generator) * 2 - 1) * bound
        if i == 0:
            w[:, mlp.n_input:] = 0
        if i == len(mlp.layer_shapes()) - 1:
            w[mlp.n_output:] = 0
        parts.append(w.flatten())
    return torch.cat(parts)


def init_params(grid: GridDesc, mlp: MlpDesc, seed: int = 1337, device="cuda") -> torch.Tensor:
    """[MLP | grid] float32 parameters: Xavier MLP, grid uniform in +-1e-4."""
    g = torch.Generator().manual_seed(seed)
    w = xavier_init(mlp, g)
    table = (torch.rand(grid.n_params, generator=g) * 2 - 1) * 1e-4
    return torch.cat([w, table]).to(device)
