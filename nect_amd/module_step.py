"""Training step for models the fused step does not cover: composite encodings (QuadCubes, HyperCubes, a grid
with Identity) and networks wider than 64 neurons.

The batch body of BaseTrainer.fit (nect/trainers/base_trainer.py:517-653) with the model as a torch module:

    sampler      nect_ray_setup + nect_ray_tiles + nect_ray_positions   sample positions [slots, 3] (z, y, x)
                                                                        in 32-sample tiles of each ray
    model        model(positions[, t])                                  hash grids in one launch per
                                                                        dimensionality, MLP fused or GEMMs
    integral     nect_integral_loss                                     line integrals, loss, dL/dmu [slots]
    backward     autograd from dL/dmu                                   MLP backward, hash-grid scatter
    optimiser    nect_grad_check + nect_adam_amp (16-bit) / nect_adam   dense Adam on the flat parameters

Slots past a ray's samples are evaluated but carry no gradient (dL/dmu is zero there).  The model's flat
parameter vector is the float32 master; in 16-bit the step hands the module its parameters in the compute dtype
from a copy Adam writes as it updates them (cast afresh when anything else changed the parameters), where a
module called outside the step casts them on every call.  In 16-bit a
step whose gradient is not finite is skipped on the device (the reference's precision="16-mixed"); the
modules scale the backward by a power of two of max |dL/dy| themselves.

Large batches of a model that names its grids' sample coordinates (``grid_sources``) follow the fused path's
schedule: the batch is taken in detector-patch order (nect_ray_order), so neighbouring rays' samples are
encoded together, and each grid's table gradient is scattered over groups of 64 neighbouring rays with the
coarse levels summed in LDS (nect_grid_backward_ray_groups_cols), replacing the module's per-sample scatter;
line integrals are returned in the caller's order.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .fused import LOSS_GIVEN, aggregation_plan, loss_code
from .sampling import ct_sampling as cs


class ModuleTrainStep:
    # the grouped scatter takes one workgroup per 64 rays x 16 samples; below this many it leaves the device idle
    GROUPS_MIN_WORKGROUPS = 2048

    def __init__(self, model: torch.nn.Module, dtype: torch.dtype, max_rays: int, max_points_per_ray: int,
                 loss: str = "L2", betas=(0.9, 0.999), eps: float = 1e-8, weight_decay: float = 0.0,
                 dynamic: bool = False):
        self.model = model
        self.params = model.net.params
        if self.params.dtype != torch.float32 or not self.params.is_cuda:
            raise ValueError("the model's parameters must be a float32 GPU tensor")
        self.device = self.params.device
        self.dtype = dtype
        self.dynamic = dynamic
        self.loss = loss_code(loss)
        self.betas, self.eps, self.weight_decay = betas, eps, weight_decay
        self.clip_grad_value: float | None = None
        self.m = torch.zeros_like(self.params.data)
        self.v = torch.zeros_like(self.params.data)
        # the step's parameter gradient: the module's backward adds into it, Adam clears it
        self._grad = torch.zeros_like(self.params.data)
        self._grad_clear = True
        self.step_count = 0
        self.amp = dtype != torch.float32
        self.amp_state = torch.zeros(_lib.AMP_STATE_BYTES // 4, dtype=torch.int32, device=self.device)
        if self.amp:
            _lib.call("nect_amp_init", _lib.ptr(self.amp_state), 1.0, _lib.stream_of(self.device))
        self.amp_growth_interval = 2000
        # table-gradient schedule: "groups" (ray groups, patch order) when the model names its grids' coordinates
        self.scatter = "groups" if getattr(model, "grid_sources", None) is not None else "module"
        self.presort = True
        self.agg_rounds: list | None = None      # per grid, per level; None: aggregation_plan()
        # the scatter's launches (nect_grid_backward_ray_groups' split): 0 one launch per grid, 1 the aggregated
        # and the direct levels as two
        self.scatter_split = 0
        # 16-bit: the parameters in the compute dtype, written by Adam as it updates them and handed to the
        # module's forward (instead of a cast of every parameter per step); valid while the parameters' version
        # counter and storage are those recorded when it was last written
        self.use_compute_copy = True
        self._compute_copy: torch.Tensor | None = None
        self._compute_key = None
        self._order_ws = torch.empty(0, dtype=torch.uint8, device=self.device)
        self._order_need: dict = {}
        self._plan_cache: dict = {}
        self._alloc(max_rays, max_points_per_ray)

    def _alloc(self, max_rays: int, S: int):
        dev = self.device
        self.max_rays, self.max_S = max_rays, S
        self.max_tiles = cs.tiles_needed(max_rays, S)
        slots = self.max_tiles * 32
        self.rays = torch.empty((max_rays, _lib.RAY_BYTES), dtype=torch.uint8, device=dev)
        self.tile_offsets = torch.empty(max_rays + 1, dtype=torch.int32, device=dev)
        self.tile_ray = torch.empty(self.max_tiles, dtype=torch.int32, device=dev)
        self.n_tiles = torch.empty(1, dtype=torch.int32, device=dev)
        self.pos = torch.empty((slots, 3), dtype=torch.float32, device=dev)
        self.valid = torch.empty(slots, dtype=torch.int32, device=dev)
        self.dmu = torch.empty(slots, dtype=torch.float32, device=dev)
        self.y_pred = torch.empty(max_rays, dtype=torch.float32, device=dev)
        self.partials = torch.empty((max_rays + 3) // 4, dtype=torch.float32, device=dev)
        self.loss_value = torch.empty(1, dtype=torch.float32, device=dev)
        self.order = torch.empty(max_rays, dtype=torch.int32, device=dev)
        self._y_sorted = torch.empty(max_rays, dtype=torch.float32, device=dev)

    def _positions(self, geometry, angle, ray_index, start, num_rays, S, step_mm, uniform, jitter, offset):
        """Rays, their tiles and the sample positions of the batch; returns the slot count."""
        if num_rays > self.max_rays or S > self.max_S:
            self._alloc(max(num_rays, self.max_rays), max(S, self.max_S))
        st = _lib.stream_of(self.device)
        rays = self.rays[:num_rays]
        _lib.call("nect_ray_setup", geometry._struct(), float(angle), _lib.ptr(ray_index), ray_index.numel(),
                  int(num_rays), int(start), int(S), float(step_mm), int(bool(uniform)), _lib.ptr(jitter),
                  _lib.ptr(offset), _lib.ptr(rays), st)
        max_tiles = cs.tiles_needed(num_rays, S)
        _lib.call("nect_ray_tiles", _lib.ptr(rays), int(num_rays), int(max_tiles), _lib.ptr(self.tile_offsets),
                  _lib.ptr(self.tile_ray), _lib.ptr(self.n_tiles), st)
        _lib.call("nect_ray_positions", _lib.ptr(rays), _lib.ptr(self.tile_offsets), _lib.ptr(self.tile_ray),
                  _lib.ptr(self.n_tiles), int(max_tiles), _lib.ptr(self.pos), _lib.ptr(self.valid), st)
        return max_tiles * 32

    def _params_key(self):
        return (self.params._version, self.params.data_ptr())

    def _compute_params(self) -> torch.Tensor | None:
        """The parameters in the compute dtype for the module's forward: the copy Adam wrote, cast afresh when
        anything else has changed the parameters since (None in float32 or when switched off)."""
        if not (self.amp and self.use_compute_copy and hasattr(self.model.net, "compute_params")):
            return None
        if self._compute_copy is None or self._compute_key != self._params_key():
            if self._compute_copy is None:
                self._compute_copy = torch.empty(self.params.shape, dtype=self.dtype, device=self.device)
            _lib.call("nect_cast", _lib.ptr(self.params.data), _lib.ptr(self._compute_copy),
                      _lib.dtype_code(self.dtype), self.params.numel(), _lib.stream_of(self.device))
            self._compute_key = self._params_key()
        return self._compute_copy

    def _mu(self, slots: int, t: float):
        x = self.pos[:slots]
        net = self.model.net
        cp = self._compute_params()
        if cp is not None:
            net.compute_params = cp
        try:
            out = self.model(x, t) if self.dynamic else self.model(x)
        finally:
            if cp is not None:
                net.compute_params = None
        return out[:, 0]

    def _integral(self, num_rays, mu, target, l_max, mode, y_out):
        self.dmu.zero_()
        _lib.call("nect_integral_loss", _lib.ptr(self.rays[:num_rays]), _lib.ptr(self.tile_offsets), int(num_rays),
                  _lib.ptr(mu), _lib.ptr(target), float(l_max), int(mode), _lib.ptr(y_out), _lib.ptr(self.dmu),
                  _lib.ptr(self.partials), _lib.ptr(self.loss_value), _lib.stream_of(self.device))

    # ------------------------------------------------------------ ray groups

    def _use_groups(self, num_rays: int, points_per_ray: int) -> bool:
        passes = 2 * ((points_per_ray + 31) // 32)
        return self.scatter == "groups" and ((num_rays + 63) // 64) * passes >= self.GROUPS_MIN_WORKGROUPS

    def _ray_order(self, geometry, ray_index: torch.Tensor, start: int, num_rays: int):
        """self.order: the batch's rays grouped by 8x8 detector patch (nect_ray_order)."""
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
        _lib.call("nect_ray_order", g, _lib.ptr(ray_index), ray_index.numel(), int(num_rays), int(start),
                  _lib.ptr(self.order), _lib.ptr(self._order_ws), self._order_ws.numel(), _lib.stream_of(self.device))

    def _sorted_index(self, geometry, ray_index: torch.Tensor, start: int, num_rays: int, perm) -> torch.Tensor:
        """ray_index[start + perm] (slots past the index get a pixel outside the detector: empty rays)."""
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

    def _plan(self, k: int, desc, geometry, step_mm: float, num_rays: int):
        if self.agg_rounds is not None:
            return (ctypes.c_int32 * desc.n_levels)(*self.agg_rounds[k])
        key = (k, tuple(geometry.nDetector), tuple(geometry.dDetector), float(step_mm), int(num_rays))
        plan = self._plan_cache.get(key)
        if plan is None:
            n_pix = geometry.nDetector[0] * geometry.nDetector[1]
            plan = (ctypes.c_int32 * desc.n_levels)(*aggregation_plan(desc, geometry, step_mm,
                                                                      num_rays / max(1, n_pix)))
            if len(self._plan_cache) > 64:
                self._plan_cache.clear()
            self._plan_cache[key] = plan
        return plan

    def _grid_scatter(self, geometry, num_rays: int, points_per_ray: int, step_mm: float, t: float, ordered: bool,
                      dfeat: torch.Tensor, grad: torch.Tensor, dfeat_scale: torch.Tensor | None = None):
        """Each grid's table gradient from dL/dfeatures [slots, n_encoded] (= dfeat * dfeat_scale when the scale
        is given: the network's 16-bit loss-scaled result, unscaled as the scatter reads it) over the batch's ray
        groups."""
        st = _lib.stream_of(self.device)
        rays = self.rays[:num_rays]
        if dfeat.stride(1) != 1:
            dfeat = dfeat.contiguous()
        if dfeat_scale is not None:
            dfeat_scale = dfeat_scale.float().contiguous()
        n_enc = dfeat.stride(0)
        esz = dfeat.element_size()
        for k, (part, src) in enumerate(zip(self.model.net.grid_parts(), self.model.grid_sources)):
            cols = (ctypes.c_int32 * len(src))(*src)
            _lib.call("nect_grid_backward_ray_groups_cols", part.desc.struct_ref(), cols, _lib.ptr(rays),
                      _lib.ptr(self.tile_offsets), int(num_rays), _lib.ptr(self.order) if ordered else None,
                      int(points_per_ray), float(t), _lib.dtype_code(dfeat.dtype), dfeat.data_ptr() + esz * part.out0,
                      n_enc, _lib.ptr(dfeat_scale), grad.data_ptr() + 4 * part.p0,
                      self._plan(k, part.desc, geometry, step_mm, num_rays), int(self.scatter_split), st)

    def _backward(self, mu: torch.Tensor, slots: int, geometry, num_rays: int, points_per_ray: int, step_mm: float,
                  t: float, groups: bool, ordered: bool):
        net = self.model.net
        if groups:
            net.encoding_backward = lambda dfeat, grad, scale=None: self._grid_scatter(
                geometry, num_rays, points_per_ray, step_mm, t, ordered, dfeat, grad, scale)
        net.grad_buffer = self._grad
        self._grad_clear = False
        try:
            mu.backward(self.dmu[:slots])
        finally:
            net.encoding_backward = None
            net.grad_buffer = None

    # ------------------------------------------------------------------ steps

    def step(self, geometry, angle: float, ray_index: torch.Tensor, starting_ray_index: int, num_rays: int,
             points_per_ray: int, step_mm: float, uniform: bool, jitter: torch.Tensor, detector_offset: torch.Tensor,
             target: torch.Tensor, l_max: float, lr: float, t: float = 0.0, optimizer_step: bool = True):
        groups = self._use_groups(num_rays, points_per_ray)
        if num_rays > self.max_rays or points_per_ray > self.max_S:
            self._alloc(max(num_rays, self.max_rays), max(points_per_ray, self.max_S))
        perm = None
        if groups and self.presort:
            self._ray_order(geometry, ray_index, starting_ray_index, num_rays)
            perm = self.order[:num_rays].long()
            ray_index = self._sorted_index(geometry, ray_index, starting_ray_index, num_rays, perm)
            starting_ray_index = 0
            jitter, detector_offset = jitter[:num_rays][perm], detector_offset[:num_rays][perm]
            target = target[:num_rays][perm]
        slots = self._positions(geometry, angle, ray_index, starting_ray_index, num_rays, points_per_ray, step_mm,
                                uniform, jitter, detector_offset)
        mu = self._mu(slots, t).contiguous()
        y_buf = self._y_sorted if perm is not None else self.y_pred
        self._integral(num_rays, mu.detach(), target.contiguous(), l_max, self.loss, y_buf)
        if perm is not None:
            self.y_pred[:num_rays].index_copy_(0, perm, y_buf[:num_rays])
        elif groups:
            self._ray_order(geometry, ray_index, starting_ray_index, num_rays)
        self._zero_grad()
        self._backward(mu, slots, geometry, num_rays, points_per_ray, step_mm, t, groups, ordered=perm is None)
        self._last_rays = num_rays
        if optimizer_step:
            self.adam(lr)
        return self.loss_value

    def step_accumulated(self, geometry, angles, ray_index, starting_ray_index, num_rays, points_per_ray, step_mm,
                         uniform, jitters, detector_offsets, target, l_max, lr, t: float = 0.0,
                         optimizer_step: bool = True):
        """Continuous-scanning step (nect/trainers/continous_scanning_trainer.py:81-167): the prediction is the
        mean of the K sub-angles' line integrals, summed as y_k / K in sub-angle order; each sub-angle is then
        recomputed with autograd and back-propagated with dL/dy / K."""
        K = len(angles)
        ys = torch.empty((K, num_rays), dtype=torch.float32, device=self.device)
        zeros = torch.zeros(num_rays, dtype=torch.float32, device=self.device)
        for k in range(K):
            slots = self._positions(geometry, angles[k], ray_index, starting_ray_index, num_rays, points_per_ray,
                                    step_mm, uniform, jitters[k], detector_offsets[k])
            with torch.no_grad():
                mu = self._mu(slots, t).contiguous()
            self._integral(num_rays, mu, zeros, l_max, LOSS_GIVEN, ys[k])
        y = ys[0] / K
        for k in range(1, K):
            y = y + ys[k] / K
        self.y_pred[:num_rays].copy_(y)
        diff = y - target[:num_rays]
        if self.loss == 1:
            loss = torch.mean(diff * diff)
            gy = 2.0 * diff / num_rays
        else:
            loss = torch.mean(diff.abs())
            gy = torch.sign(diff) / num_rays
        gy = (gy / K).contiguous()
        self._zero_grad()
        groups = self._use_groups(num_rays, points_per_ray)
        if groups:
            self._ray_order(geometry, ray_index, starting_ray_index, num_rays)   # the same pixels at every sub-angle
        for k in range(K):
            slots = self._positions(geometry, angles[k], ray_index, starting_ray_index, num_rays, points_per_ray,
                                    step_mm, uniform, jitters[k], detector_offsets[k])
            mu = self._mu(slots, t).contiguous()
            self._integral(num_rays, mu.detach(), gy, l_max, LOSS_GIVEN, ys[k])
            self._backward(mu, slots, geometry, num_rays, points_per_ray, step_mm, t, groups, ordered=True)
        self.loss_value.copy_(loss.reshape(1))
        self._last_rays = num_rays
        if optimizer_step:
            self.adam(lr)
        return self.loss_value

    @property
    def grad(self) -> torch.Tensor:
        """The last step's parameter gradient (zero after the optimiser step)."""
        return self._grad

    def _zero_grad(self):
        # Adam leaves the gradient zero; only a step without one left it to clear
        if not self._grad_clear:
            self._grad.zero_()
        self._grad_clear = True

    def adam(self, lr: float):
        self.step_count += 1
        st = _lib.stream_of(self.device)
        p, g = self.params.data, self.grad
        self._grad_clear = True   # both kernels below leave it zero
        clip = float(self.clip_grad_value or 0.0)
        if self.amp:
            # the compute copy is updated with the parameters (both, or on a skipped step neither) when it is
            # current; a stale one is left to the next forward's cast
            copy = self._compute_copy if (self._compute_copy is not None and
                                          self._compute_key == self._params_key()) else None
            _lib.call("nect_grad_check", _lib.ptr(g), g.numel(), _lib.ptr(self.amp_state), st)
            _lib.call("nect_adam_amp", _lib.ptr(p), _lib.ptr(g), _lib.ptr(self.m), _lib.ptr(self.v), _lib.ptr(copy),
                      _lib.dtype_code(self.dtype) if copy is not None else _lib.F32, p.numel(), float(lr),
                      float(self.betas[0]), float(self.betas[1]), float(self.eps), float(self.weight_decay),
                      _lib.ptr(self.amp_state), 2.0, 0.5, int(self.amp_growth_interval), clip, st)
            if copy is not None:
                self._compute_key = self._params_key()   # the kernel's writes leave the version counter alone
            return
        _lib.call("nect_adam", _lib.ptr(p), _lib.ptr(g), _lib.ptr(self.m), _lib.ptr(self.v), None, _lib.F32,
                  p.numel(), float(lr), float(self.betas[0]), float(self.betas[1]), float(self.eps),
                  float(self.weight_decay), self.step_count, clip, st)

    def amp_status(self) -> dict:
        st = self.amp_state.cpu()
        return {"scale": float(st[:1].view(torch.float32).item()), "step": int(st[2]), "skipped": int(st[4])}

    def adam_steps(self) -> int:
        return self.amp_status()["step"] if self.amp else self.step_count

    def set_adam_steps(self, n: int):
        self.step_count = int(n)
        if self.amp:
            self.amp_state[2] = int(n)

    def refresh_shadow(self):
        """After the float32 parameters were written outside the step: the next forward casts them afresh (the
        version check sees writes made through the parameter itself; this also covers writes through .data)."""
        self._compute_key = None

    def last_samples(self) -> int:
        _, _, count, _ = cs.ray_fields(self.rays[: self._last_rays])
        return int(count.sum().item())
