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
parameter vector is the float32 master; the modules cast it to the compute dtype on every call.  In 16-bit a
step whose gradient is not finite is skipped on the device (the reference's precision="16-mixed"); the
modules scale the backward by a power of two of max |dL/dy| themselves.
"""
from __future__ import annotations

import torch

from . import _lib
from .fused import LOSS_GIVEN, loss_code
from .sampling import ct_sampling as cs


class ModuleTrainStep:
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
        self.step_count = 0
        self.amp = dtype != torch.float32
        self.amp_state = torch.zeros(_lib.AMP_STATE_BYTES // 4, dtype=torch.int32, device=self.device)
        if self.amp:
            _lib.call("nect_amp_init", _lib.ptr(self.amp_state), 1.0, _lib.stream_of(self.device))
        self.amp_growth_interval = 2000
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

    def _mu(self, slots: int, t: float):
        x = self.pos[:slots]
        out = self.model(x, t) if self.dynamic else self.model(x)
        return out[:, 0]

    def _integral(self, num_rays, mu, target, l_max, mode, y_out):
        self.dmu.zero_()
        _lib.call("nect_integral_loss", _lib.ptr(self.rays[:num_rays]), _lib.ptr(self.tile_offsets), int(num_rays),
                  _lib.ptr(mu), _lib.ptr(target), float(l_max), int(mode), _lib.ptr(y_out), _lib.ptr(self.dmu),
                  _lib.ptr(self.partials), _lib.ptr(self.loss_value), _lib.stream_of(self.device))

    def step(self, geometry, angle: float, ray_index: torch.Tensor, starting_ray_index: int, num_rays: int,
             points_per_ray: int, step_mm: float, uniform: bool, jitter: torch.Tensor, detector_offset: torch.Tensor,
             target: torch.Tensor, l_max: float, lr: float, t: float = 0.0, optimizer_step: bool = True):
        slots = self._positions(geometry, angle, ray_index, starting_ray_index, num_rays, points_per_ray, step_mm,
                                uniform, jitter, detector_offset)
        mu = self._mu(slots, t).contiguous()
        self._integral(num_rays, mu.detach(), target.contiguous(), l_max, self.loss, self.y_pred)
        self.params.grad = None
        mu.backward(self.dmu[:slots])
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
        self.params.grad = None
        for k in range(K):
            slots = self._positions(geometry, angles[k], ray_index, starting_ray_index, num_rays, points_per_ray,
                                    step_mm, uniform, jitters[k], detector_offsets[k])
            mu = self._mu(slots, t).contiguous()
            self._integral(num_rays, mu.detach(), gy, l_max, LOSS_GIVEN, ys[k])
            mu.backward(self.dmu[:slots])
        self.loss_value.copy_(loss.reshape(1))
        self._last_rays = num_rays
        if optimizer_step:
            self.adam(lr)
        return self.loss_value

    @property
    def grad(self) -> torch.Tensor:
        if self.params.grad is None:
            self.params.grad = torch.zeros_like(self.params.data)
        return self.params.grad

    def adam(self, lr: float):
        self.step_count += 1
        st = _lib.stream_of(self.device)
        p, g = self.params.data, self.grad
        clip = float(self.clip_grad_value or 0.0)
        if self.amp:
            _lib.call("nect_grad_check", _lib.ptr(g), g.numel(), _lib.ptr(self.amp_state), st)
            _lib.call("nect_adam_amp", _lib.ptr(p), _lib.ptr(g), _lib.ptr(self.m), _lib.ptr(self.v), None,
                      _lib.F32, p.numel(), float(lr), float(self.betas[0]), float(self.betas[1]), float(self.eps),
                      float(self.weight_decay), _lib.ptr(self.amp_state), 2.0, 0.5, int(self.amp_growth_interval),
                      clip, st)
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
        """The modules cast the float32 parameters on every call: nothing to refresh."""

    def last_samples(self) -> int:
        _, _, count, _ = cs.ray_fields(self.rays[: self._last_rays])
        return int(count.sum().item())
