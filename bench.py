"""Training-throughput benchmark of the fused HIP path (one MI355X).

Workload (the 256^3 cone-beam configuration): detector 256x256 at 1 mm, SOD
400 mm, SDD 800 mm, 0.5 mm voxels (128 mm field), 400 views over 360 deg;
hash grid 16 levels x 2 features, table 2^19, base resolution 16, growth 1.5;
MLP 2x64 ReLU with Softplus output; 2^16 rays per batch (one view), up to
256 samples per ray (uniform spacing L_max/256); L2 loss; Adam.  Precision:
float16 table, features and MLP (float32 accumulation, float32 line integral,
loss, gradients and Adam state).

A step = ray setup -> encoder -> MLP -> line integral + loss -> MLP backward
-> grid scatter -> Adam, over one batch.  Data: analytic float64 projections
of a seeded packed-sphere phantom for 8 views, cycled (synthetic).

Prints one JSON line.  Extra fields: ``roofline`` of the dominant kernel
(per-launch duration from HIP events recorded around that launch inside the
timed region), ``kernels`` (mean ms of every kernel), ``fp32_torch`` (the
plain float32 PyTorch-eager GPU path on the same batches) and ``cpu_baseline``
(the CPU oracle's training step on a bounded sample, host cores).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dtype", default="float16", choices=["float16", "bfloat16", "float32"])
    ap.add_argument("--no-fp32-torch", action="store_true", help="skip timing the plain float32 torch path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="HBM bytes per launch measured by rocprofv3 counters (optional)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (to scale FETCH_SIZE/WRITE_SIZE counters)")
    return ap.parse_args()


def copy_calibration():
    """A plain device-to-device copy of known size: 1 GiB read + 1 GiB written per launch, 5 launches."""
    dev = torch.device("cuda", 0)
    src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": src.numel() * 4, "launches": 5}), flush=True)


class Workload:
    """The 256^3 cone-beam training configuration."""

    def __init__(self, device, dtype, n_views_data=8):
        from nect_amd.sampling.geometry import Geometry
        from nect_amd import phantom

        self.device = device
        self.dtype = dtype
        self.n_angles = 400
        self.geometry = Geometry(nDetector=(256, 256), dDetector=(1.0, 1.0), mode="cone", DSD=800.0, DSO=400.0,
                                 nVoxel=(256, 256, 256), dVoxel=(0.5, 0.5, 0.5),
                                 angles=[2 * math.pi * i / 400 for i in range(400)])
        self.c_geometry = self.geometry.get_c_geometry()
        self.S = 256
        self.B = 256 * 256
        self.l_max = self.geometry.max_distance_traveled
        self.step_mm = self.l_max / self.S
        ph = phantom.packed_spheres(256, 0.5, seed=1, max_spheres=1200)
        self.view_ids = [int(i * self.n_angles / n_views_data) for i in range(n_views_data)]
        angles = [self.geometry.angles[i] for i in self.view_ids]
        proj = phantom.project(ph, "cone", (256, 256), (1.0, 1.0), 400.0, 800.0, angles, device=device)
        self.proj = (proj / proj.max()).reshape(len(angles), -1).contiguous()
        self.gen = torch.Generator(device=device).manual_seed(0)

    def batch(self, k: int):
        """(angle, ray_index, jitter, detector_offset, target) of step k."""
        v = k % len(self.view_ids)
        angle = float(self.geometry.angles[self.view_ids[v]])
        perm = torch.randperm(self.B, device=self.device, generator=self.gen)
        jitter = torch.rand(self.B, device=self.device, generator=self.gen) - 0.5
        offset = torch.zeros((self.B, 2), device=self.device)
        target = self.proj[v][perm].contiguous()
        return angle, perm, jitter, offset, target


def fused_step_factory(w: Workload, profile: bool):
    from nect_amd.fused import FusedTrainStep, init_params
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc

    grid = GridDesc(3, 16, 2, 19, 16, 1.5)
    mlp = MlpDesc(32, 1, 64, 2, "ReLU", "Softplus")
    params = init_params(grid, mlp, seed=1337, device=w.device)
    fs = FusedTrainStep(grid, mlp, params, w.dtype, max_rays=w.B, max_points_per_ray=w.S, loss="L2")
    fs.profile = profile
    return fs, grid, mlp


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see DESIGN.md for how the training "
              "step would shard). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    dtype = {"float16": torch.float16, "bfloat16": torch.bfloat16, "float32": torch.float32}[args.dtype]
    print(f"[bench] building workload", file=sys.stderr, flush=True)
    w = Workload(dev, dtype)
    fs, grid, mlp = fused_step_factory(w, profile=True)
    lr = 1e-2

    def run(k):
        angle, perm, jitter, offset, target = w.batch(k)
        return fs.step(w.c_geometry, angle, perm, 0, w.B, w.S, w.step_mm, True, jitter, offset, target, w.l_max,
                       lr=lr)

    print(f"[bench] warmup {args.warmup}", file=sys.stderr, flush=True)
    for k in range(args.warmup):
        run(k)
    torch.cuda.synchronize()
    fs.reset_timers()
    batches = [w.batch(args.warmup + k) for k in range(args.steps)]
    torch.cuda.synchronize()
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for k in range(args.steps):
        angle, perm, jitter, offset, target = batches[k]
        fs.step(w.c_geometry, angle, perm, 0, w.B, w.S, w.step_mm, True, jitter, offset, target, w.l_max, lr=lr)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    value = w.B / (ms / 1e3)
    kernels = fs.timer_means()
    n_samples = int(fs.last_samples())
    final_loss = float(fs.loss_value.item())

    # roofline of the dominant kernel: algorithmic bytes per launch / mean launch time
    per_sample_bytes = {
        "grid_backward": 12 + 16 * 2 * 4 + 8 * 16 * 2 * 4,        # coords + dfeat + atomic payload
        "grid_encode": 12 + 8 * 16 * 2 * (2 if dtype != torch.float32 else 4) + 16 * 2 * (2 if dtype != torch.float32 else 4),
    }
    dominant = max((k for k in kernels if k in per_sample_bytes), key=lambda k: kernels[k])
    achieved = per_sample_bytes[dominant] * n_samples / (kernels[dominant] / 1e3) / 1e9
    traffic = None
    try:
        with open(args.traffic_file) as f:
            traffic = json.load(f).get(dominant)
    except (OSError, ValueError):
        traffic = None
    roofline = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes_per_launch": per_sample_bytes[dominant] * n_samples,
                "samples_per_launch": n_samples}

    out = {
        "metric": "train rays/s fwd+bwd (FP16 vs FP32) at equal PSNR; export query p50/p99 ms",
        "value": round(value, 1),
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {torch.float16: "fp16", torch.bfloat16: "bf16", torch.float32: "fp32"}[dtype],
        "data": "synthetic: analytic projections of a seeded packed-sphere phantom, 8 views cycled",
        "config": {"workload": "static cone-beam 256^3, detector 256x256, 400 views, 2^16 rays/batch, "
                               "<=256 samples/ray, hash grid 16x2 T=2^19, MLP 2x64, L2, Adam",
                   "global_batch": w.B, "samples_per_ray_max": w.S, "parallelism": "single GPU"},
        "wall_s_timed": round(wall, 3),
        "samples_per_step": n_samples,
        "final_loss": final_loss,
        "kernels": {k: round(v, 4) for k, v in kernels.items()},
        "roofline": roofline,
    }

    if not args.no_fp32_torch:
        out["fp32_torch"] = time_fp32_torch(w, grid, mlp)
        out["speedup_vs_fp32_torch"] = round(value / out["fp32_torch"]["value"], 2)
    if not args.no_cpu:
        out["cpu_baseline"] = time_cpu_oracle(w, grid, mlp, args.cpu_seconds)
    print(json.dumps(out), flush=True)


def time_fp32_torch(w: Workload, grid, mlp, steps: int = 3):
    """Plain float32 PyTorch-eager GPU path on the same batches (same parameters at start)."""
    from nect_amd.baseline_torch import TorchTrainStep
    from nect_amd.fused import init_params

    print("[bench] timing plain fp32 torch path", file=sys.stderr, flush=True)
    ts = TorchTrainStep(grid, mlp, init_params(grid, mlp, seed=1337, device=w.device), lr=1e-2, loss="L2")
    b = [w.batch(1000 + k) for k in range(steps + 1)]
    angle, perm, jitter, offset, target = b[0]
    ts.step(w.c_geometry, angle, perm, 0, w.B, w.S, w.step_mm, True, jitter, offset, target, w.l_max)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        angle, perm, jitter, offset, target = b[k + 1]
        ts.step(w.c_geometry, angle, perm, 0, w.B, w.S, w.step_mm, True, jitter, offset, target, w.l_max)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    return {"value": round(w.B / (ms / 1e3), 1), "unit": "rays/s", "ms_per_step": round(ms, 2), "steps": steps,
            "dtype": "fp32", "kind": "torch-eager"}


def time_cpu_oracle(w: Workload, grid, mlp, seconds: float):
    """The oracle's training step (C sampler + numpy grid + float64 torch MLP/autograd) on host cores."""
    import numpy as np

    from oracle import hashgrid as hg
    from oracle.train import train_step
    from nect_amd.fused import init_params

    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    spec = hg.GridSpec(3, 16, 2, 19, 16, 1.5)
    cfg = {"n_params": mlp.n_params, "n_input_padded": 32, "n_hidden_layers": 2, "n_output_padded": 16,
           "activation": "ReLU", "output_activation": "Softplus"}
    params = init_params(grid, mlp, seed=1337, device="cpu").double().numpy()
    g = w.geometry
    from oracle.geometry import OracleGeometry

    og = OracleGeometry(mode="cone", nDetector=(256, 256), dDetector=(1.0, 1.0), DSD=800.0, DSO=400.0,
                        nVoxel=(256, 256, 256), dVoxel=(0.5, 0.5, 0.5))
    fields = og.c_struct_fields()
    B = 256
    rng = np.random.default_rng(0)
    print("[bench] timing CPU oracle", file=sys.stderr, flush=True)
    rays = 0
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        idx = rng.permutation(256 * 256).astype(np.int64)
        jit = (rng.random(B) - 0.5).astype(np.float32)
        tgt = rng.random(B)
        train_step(fields, float(g.angles[steps % 400]), idx, 0, B, 256, w.step_mm, True, jit,
                   np.zeros((B, 2), np.float32), tgt, w.l_max, spec, cfg, params)
        rays += B
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": round(rays / dt, 1), "unit": "rays/s", "cores": torch.get_num_threads(), "kind": "port",
            "sample": f"{steps} oracle training steps of {B} rays x <=256 samples (same grid/MLP config), "
                      f"{dt:.1f} s, no Adam"}


if __name__ == "__main__":
    main()
