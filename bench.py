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
MFMA_PEAK_TFLOPS = {"fp16": 2500.0, "bf16": 2500.0, "fp32": 157.3}
# query work per voxel of the 4-D serving model: 16 levels x 16 corners x 2 bf16 features gathered + a 2-byte
# output; MLP 32 -> 64 -> 64 -> 64 -> 1 forward (the output row alone)
QUERY_BYTES_PER_VOXEL = 16 * 16 * 2 * 2 + 2
QUERY_FLOP_PER_VOXEL = 2 * (32 * 64 + 2 * 64 * 64 + 64 * 1)


I0_CONFIG2 = 1e4          # photons per pixel of the 256^3 workload (Poisson noise, then -log)
PMAX_CONFIG2 = 4.0        # largest noiseless line integral of view 0 (sets the attenuation scale)
I0_CONFIG3 = 5e3          # photons per pixel of the 4D workload
PMAX_CONFIG3 = 4.0
_PHANTOM_CACHE: dict = {}
_TRAINED_4D: dict = {}    # parameters of the 4D model trained by train_4d_bf16 (served by serving_bench)


def config2_phantom(dev):
    """The 256^3 workload's phantom: grains of radius 4-12 voxels packed to porosity ~0.35 (attenuation 1.0),
    pore fluid 0.05, cylinder wall 0.2 (seed 1), every attenuation scaled so that the largest line integral of
    view 0 is PMAX_CONFIG2 (transmission e^-4 at the darkest pixel).  Returns (phantom, porosity)."""
    from nect_amd import phantom

    key = str(dev)
    if key not in _PHANTOM_CACHE:
        ph = phantom.relaxed_packed_spheres(256, 0.5, seed=1, device=dev)
        p0 = phantom.project(ph, "cone", (256, 256), (1.0, 1.0), 400.0, 800.0, [0.0], device=dev)
        ph = phantom.scaled(ph, PMAX_CONFIG2 / float(p0.max()))
        _PHANTOM_CACHE[key] = (ph, phantom.union_porosity(ph, 256, 0.5, device=dev))
    return _PHANTOM_CACHE[key]


def config3_phantom(dev):
    """The 4D workload's sample: the 256^3 packing (seed 3) with a brine front (attenuation 0.3) rising from
    the bottom of the volume at t = 0 to its top at t = 1, 20 % of the grains soluble (seed 4); attenuations
    scaled so that the largest line integral of view 0 at t = 0 is PMAX_CONFIG3."""
    from nect_amd import phantom

    key = ("c3", str(dev))
    if key not in _PHANTOM_CACHE:
        base = phantom.relaxed_packed_spheres(256, 0.5, seed=3, device=dev)
        ph = phantom.brine_front(base, 256, 0.5, soluble_fraction=0.2, brine=0.3, seed=4)
        p0 = phantom.project(ph, "cone", (512, 512), (0.5, 0.5), 400.0, 800.0, [0.0], device=dev, times=[0.0])
        _PHANTOM_CACHE[key] = phantom.scaled(ph, PMAX_CONFIG3 / float(p0.max()))
    return _PHANTOM_CACHE[key]


class Views4D:
    """Noisy projections of the 4D sample for the views a run uses, made on first use: view v (of 2000 over
    5 turns) at t = v / 1999, Poisson noise at I0_CONFIG3 (seeded by v) then -log, divided by PMAX_CONFIG3."""

    def __init__(self, dev, geo, n_views):
        self.dev, self.geo, self.n_views = dev, geo, n_views
        self.ph = config3_phantom(dev)
        self.cache = {}

    def __call__(self, v: int) -> torch.Tensor:
        from nect_amd import phantom

        if v not in self.cache:
            t = v / (self.n_views - 1)
            p = phantom.project(self.ph, "cone", (512, 512), (0.5, 0.5), 400.0, 800.0, [self.geo.angles[v]],
                                device=self.dev, times=[t])
            g = torch.Generator(device=self.dev).manual_seed(1000 + v)
            self.cache[v] = (phantom.poisson_log(p, I0_CONFIG3, generator=g) / PMAX_CONFIG3).reshape(-1)
        return self.cache[v]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dtype", default="float16", choices=["float16", "bfloat16", "float32"])
    ap.add_argument("--no-fp32-torch", action="store_true", help="skip timing the plain float32 torch path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-serving", action="store_true", help="skip the export-query serving measurement")
    ap.add_argument("--no-psnr", action="store_true", help="skip the FP16-vs-FP32 reconstruction PSNR comparison")
    ap.add_argument("--psnr-only", action="store_true", help="only the FP16-vs-FP32 reconstruction PSNR comparison")
    ap.add_argument("--psnr-steps", type=int, default=2000)
    ap.add_argument("--psnr256-steps", type=int, default=150,
                    help="steps of the FP16-vs-FP32 comparison on the 256^3 noisy packed-sphere workload")
    ap.add_argument("--no-psnr256", action="store_true", help="skip the 256^3 FP16-vs-FP32 PSNR comparison")
    ap.add_argument("--psnr256-only", action="store_true", help="only the 256^3 FP16-vs-FP32 PSNR comparison")
    ap.add_argument("--serving-only", action="store_true", help="only the export-query serving measurement")
    ap.add_argument("--serving-seconds", type=float, default=8.0, help="arrival window per request rate")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="HBM bytes per launch measured by rocprofv3 counters (optional)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (to scale FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--atomic-calibration", action="store_true",
                    help="only measure the device's float-atomic request rate (nect_calibrate_atomics)")
    ap.add_argument("--atomic-table-mib", type=int, nargs="+", default=[256], metavar="MIB",
                    help="with --atomic-calibration: table sizes to spread the requests over (one JSON line each)")
    ap.add_argument("--quadcubes-points-per-batch", type=int, default=5_000_000,
                    help="with --quadcubes-only: the configuration's points_per_batch (half of it per step, as the "
                         "reference's Projector takes)")
    ap.add_argument("--static-preset-only", action="store_true",
                    help="only the training step of the reference's default static model (hash_grid preset)")
    ap.add_argument("--scatter-split", action="store_true",
                    help="with --quadcubes-only: the scatter's LDS-aggregated and direct levels as two launches each "
                         "(to time the phases apart)")
    ap.add_argument("--atomic-file", default=os.path.join(ROOT, "profiles", "atomic_rate.json"),
                    help="the atomic request rate measured by --atomic-calibration (optional)")
    ap.add_argument("--query-only", action="store_true",
                    help="only the 512^3 export volume of the serving model (e.g. under a profiler)")
    ap.add_argument("--summarize-pmc", nargs="+", metavar="CSV",
                    help="print the mean per-dispatch value of each counter per kernel from rocprofv3 "
                         "counter_collection.csv files, then exit (no GPU use)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the recompute-vs-store comparison and the 4D BF16 training throughput")
    ap.add_argument("--extras-only", action="store_true", help="only those two measurements")
    ap.add_argument("--quadcubes-only", action="store_true",
                    help="only the training step of the reference's default dynamic model (QuadCubes preset)")
    ap.add_argument("--train-4d-only", action="store_true",
                    help="only the 4D BF16 training-step measurement (then serving of the trained model, unless "
                         "--no-serving)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="short run for kernel profilers: uniform random targets instead of the phantom's "
                         "projections (the kernels' work depends on the geometry only), no baselines")
    return ap.parse_args()


def copy_calibration():
    """A plain device-to-device copy of known size: 1 GiB read + 1 GiB written per launch, 5 launches (the
    counters' scale), then the rate of 4 GiB copies (the device's streaming rate for a half-read, half-written
    mix, against which Adam's is judged)."""
    dev = torch.device("cuda", 0)
    src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": src.numel() * 4, "launches": 5}), flush=True)
    del src, dst
    src = torch.ones(1 << 30, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    print(json.dumps({"copy_GiB_each_way": 4, "ms": round(ms, 3),
                      "TB_per_s_read_plus_written": round(2 * src.numel() * 4 / ms / 1e9, 3)}), flush=True)


def atomic_calibration(table_mib: int = 256):
    """Float-atomic request rate: 2^22 threads x 32 adds into a table of 64-byte blocks (256 MiB: the size of
    the Infinity Cache), lanes grouped 1..16 to a block (64..4 requests per wave instruction); best of 5 launches
    per grouping."""
    from nect_amd import _lib

    dev = torch.device("cuda", 0)
    n_blocks, threads, per = table_mib << 14, 1 << 22, 32
    table = torch.zeros(n_blocks * 16, dtype=torch.float32, device=dev)
    st = _lib.stream_of(dev)
    out = {"table_bytes": n_blocks * 64, "threads": threads, "adds_per_thread": per, "lanes_per_block": {}}
    for lpb in (1, 2, 4, 8, 16):
        best = float("inf")
        for rep in range(6):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            _lib.call("nect_calibrate_atomics", _lib.ptr(table), n_blocks, lpb, per, threads, 1234 + rep, st)
            e1.record()
            torch.cuda.synchronize()
            if rep:
                best = min(best, e0.elapsed_time(e1))
        req = threads * per / lpb
        out["lanes_per_block"][str(lpb)] = {"ms": round(best, 4), "G_requests_per_s": round(req / best / 1e6, 2),
                                            "G_adds_per_s": round(threads * per / best / 1e6, 2)}
    out["requests_per_s_G"] = out["lanes_per_block"]["1"]["G_requests_per_s"]
    assert float(table.sum(dtype=torch.float64)) == 6 * 5 * threads * per   # every add landed
    print(json.dumps(out), flush=True)


class Workload:
    """The 256^3 cone-beam training configuration."""

    def __init__(self, device, dtype, n_views_data=8, phantom_targets=True):
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
        self.view_ids = [int(i * self.n_angles / n_views_data) for i in range(n_views_data)]
        angles = [self.geometry.angles[i] for i in self.view_ids]
        self.gen = torch.Generator(device=device).manual_seed(0)
        if phantom_targets:
            ph, self.porosity = config2_phantom(device)
            proj = phantom.project(ph, "cone", (256, 256), (1.0, 1.0), 400.0, 800.0, angles, device=device)
            noisy = phantom.poisson_log(proj, I0_CONFIG2, generator=torch.Generator(device=device).manual_seed(5))
            self.proj = (noisy / noisy.max()).reshape(len(angles), -1).contiguous()
        else:
            self.proj = torch.rand((len(angles), self.B), device=device, generator=self.gen)

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


def summarize_pmc(paths):
    """Mean per-dispatch counter value per (kernel, counter) over rocprofv3 counter_collection.csv files or
    rocpd databases (results.db); for a database without counters, the kernel trace (calls, mean ns)."""
    import csv
    import sqlite3
    from collections import defaultdict

    per_dispatch = defaultdict(float)
    names = {}
    trace = {}
    for path in paths:
        if path.endswith(".db"):
            db = sqlite3.connect(path)
            rows = db.execute("select dispatch_id, counter_name, value, kernel_name from counters_collection")
            n = 0
            for disp, counter, value, kname in rows:
                key = (path, str(disp), counter)
                per_dispatch[key] += float(value)
                names[key] = kname
                n += 1
            if n == 0:
                for name, calls, total, avg, pct in db.execute("select * from top_kernels"):
                    trace[name.split("(")[0][:120]] = {"calls": calls, "total_us": total, "mean_us": avg,
                                                       "percent": pct}
            continue
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                key = (path, row["Dispatch_Id"], row["Counter_Name"])
                per_dispatch[key] += float(row["Counter_Value"])
                names[key] = row["Kernel_Name"]
    if trace:
        print(json.dumps(trace, indent=1))
        if not per_dispatch:
            return
    acc = defaultdict(list)
    for key, v in per_dispatch.items():
        acc[(names[key], key[2])].append(v)
    out = {}
    for (kernel, counter), vals in sorted(acc.items()):
        short = kernel.split("(")[0][:120]
        out.setdefault(short, {})[counter] = {"mean": sum(vals) / len(vals), "dispatches": len(vals)}
    print(json.dumps(out, indent=1))


def main():
    args = parse()
    if args.summarize_pmc:
        summarize_pmc(args.summarize_pmc)
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.atomic_calibration:
        for mib in args.atomic_table_mib:
            atomic_calibration(mib)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (rays and views would shard data-parallel "
              "with one gradient exchange per step). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.serving_only:
        print(json.dumps({"serving": serving_bench(dev, args.serving_seconds, traffic_file=args.traffic_file)}),
              flush=True)
        return
    if args.query_only:
        srv = serving_model(dev, timing=True)
        print(json.dumps({"query_volume": query_volume(srv, reps=3, traffic_file=args.traffic_file)}), flush=True)
        return
    if args.psnr_only:
        print(json.dumps({"psnr": psnr_bench(dev, args.psnr_steps)}), flush=True)
        return
    if args.psnr256_only:
        print(json.dumps({"psnr_256": psnr_bench_256(dev, args.psnr256_steps)}), flush=True)
        return
    if args.quadcubes_only:
        print(json.dumps({"quadcubes": quadcubes_step(dev, split=args.scatter_split,
                                                      points_per_batch=args.quadcubes_points_per_batch)}), flush=True)
        return
    if args.train_4d_only:
        out = {"train_4d_bf16": train_4d_bf16(dev)}
        if not args.no_serving:
            out["serving"] = serving_bench(dev, args.serving_seconds, traffic_file=args.traffic_file)
        print(json.dumps(out), flush=True)
        return
    if args.extras_only:
        print(json.dumps({"recompute_vs_store": recompute_vs_store(dev), "train_4d_bf16": train_4d_bf16(dev),
                          "quadcubes": quadcubes_step(dev)}), flush=True)
        return
    dtype = {"float16": torch.float16, "bfloat16": torch.bfloat16, "float32": torch.float32}[args.dtype]
    print(f"[bench] building workload", file=sys.stderr, flush=True)
    w = Workload(dev, dtype, phantom_targets=not args.kernels_only)
    if args.static_preset_only:
        print(json.dumps({"static_preset": static_preset_step(dev, w)}), flush=True)
        return
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
    b = 4 if dtype == torch.float32 else 2
    L, F, D = grid.n_levels, grid.n_features, grid.n_dims
    per_sample_bytes = {
        "grid_backward": 4 * D + L * F * 4 + (1 << D) * L * F * 4,   # coords + dfeat + atomic payload
        "grid_encode": 4 * D + (1 << D) * L * F * b + L * F * b,     # coords + gathers + features
    }
    dominant = max((k for k in kernels if k in per_sample_bytes), key=lambda k: kernels[k])
    achieved = per_sample_bytes[dominant] * n_samples / (kernels[dominant] / 1e3) / 1e9
    measured = {}
    try:
        with open(args.traffic_file) as f:
            measured = json.load(f)
    except (OSError, ValueError):
        measured = {}
    traffic = measured.get(dominant)
    roofline = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes_per_launch": per_sample_bytes[dominant] * n_samples,
                "samples_per_launch": n_samples}
    if dominant == "grid_backward" and measured.get("grid_backward_atomic_requests"):
        # the bound this kernel meets: atomic requests (one per 64-byte block per wave instruction), counted by
        # TCC_EA0_ATOMIC_sum on the same workload, against the request rate --atomic-calibration measured
        req = measured["grid_backward_atomic_requests"]
        rate = req / (kernels[dominant] / 1e3) / 1e9
        try:
            with open(args.atomic_file) as f:
                peak = float(json.load(f)["requests_per_s_G"])
        except (OSError, ValueError, KeyError):
            peak = None
        roofline["atomic_requests"] = {"per_launch": req, "achieved": round(rate, 2), "peak": peak,
                                       "unit": "G requests/s",
                                       "frac": round(rate / peak, 3) if peak else None}

    # every kernel of the step against its own roofline (algorithmic work / mean launch time)
    H = mlp.n_hidden_layers
    mlp_fwd_flop = 2 * (mlp.n_input * mlp.width + (H - 1) * mlp.width * mlp.width + mlp.width * mlp.n_output)
    mfma_peak = 157.3 if dtype == torch.float32 else 2500.0
    n_params = fs.params.numel()
    work = {
        "ray_setup": (84 * w.B, "GB/s"),
        "ray_order": (24 * w.B, "GB/s"),
        "grid_encode": (per_sample_bytes["grid_encode"] * n_samples, "GB/s"),
        "mlp_forward": (mlp_fwd_flop * n_samples, "TFLOP/s"),
        "integral_loss": (8 * n_samples + 16 * w.B, "GB/s"),
        "mlp_backward": (2 * mlp_fwd_flop * n_samples, "TFLOP/s"),
        "grid_backward": (per_sample_bytes["grid_backward"] * n_samples, "GB/s"),
        "adam": ((16 + 12 + b) * n_params, "GB/s"),
    }
    kernel_rooflines = {}
    kernel_rooflines["sample_points"] = sampler_roofline(w)
    for k, (amount, unit) in work.items():
        if k not in kernels or kernels[k] <= 0:
            continue
        rate = amount / (kernels[k] / 1e3) / (1e9 if unit == "GB/s" else 1e12)
        peak = HBM_PEAK_GBS if unit == "GB/s" else mfma_peak
        kernel_rooflines[k] = {"ms": round(kernels[k], 4), "achieved": round(rate, 2), "unit": unit,
                               "frac": round(rate / peak, 4)}

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
        "data": "synthetic: analytic projections of a seeded 256^3 packed-sphere phantom (porosity "
                f"{getattr(w, 'porosity', float('nan')):.3f}), Poisson noise at I0=1e4 then -log, 8 views cycled",
        "config": {"workload": "static cone-beam 256^3, detector 256x256, 400 views, 2^16 rays/batch, "
                               "<=256 samples/ray, hash grid 16x2 T=2^19, MLP 2x64, L2, Adam",
                   "global_batch": w.B, "samples_per_ray_max": w.S, "parallelism": "single GPU"},
        "wall_s_timed": round(wall, 3),
        "samples_per_step": n_samples,
        "final_loss": final_loss,
        "kernels": {k: round(v, 4) for k, v in kernels.items()},
        "rays_per_s_excluding_adam": round(w.B / ((ms - kernels.get("adam", 0.0)) / 1e3), 1),
        "kernel_rooflines": kernel_rooflines,
        "roofline": roofline,
    }

    if args.kernels_only:
        out["data"] = "synthetic: uniform random targets (kernel profiling run)"
        print(json.dumps(out), flush=True)
        return
    if not args.no_fp32_torch:
        out["fp32_torch"] = time_fp32_torch(w, grid, mlp)
        out["speedup_vs_fp32_torch"] = round(value / out["fp32_torch"]["value"], 2)
    if not args.no_psnr:
        out["psnr"] = psnr_bench(dev, args.psnr_steps)
    if not args.no_psnr256:
        out["psnr_256"] = psnr_bench_256(dev, args.psnr256_steps)
    if not args.no_extras:
        del fs
        torch.cuda.empty_cache()
        fs = None
        out["static_preset"] = static_preset_step(dev, w)
        torch.cuda.empty_cache()
        out["recompute_vs_store"] = recompute_vs_store(dev)
        out["train_4d_bf16"] = train_4d_bf16(dev)
        torch.cuda.empty_cache()
        out["quadcubes"] = quadcubes_step(dev)
        torch.cuda.empty_cache()
    if not args.no_serving:
        del fs
        torch.cuda.empty_cache()
        out["serving"] = serving_bench(dev, args.serving_seconds, traffic_file=args.traffic_file)
    if not args.no_cpu:
        out["cpu_baseline"] = time_cpu_baseline(w, grid, mlp, args.cpu_seconds)
    print(json.dumps(out), flush=True)


def sampler_roofline(w: "Workload", reps: int = 10) -> dict:
    """The drop-in sampler (nect_sample, the [B, S, 3] points of ct_sampling.sample) on the workload's batch:
    mean launch time from HIP events on the launch stream, against 20 B read + 12*S B + 8 B written per ray
    (the [B, S, 3] output is written whole, zeros past each ray's count, as the reference's interface returns
    it)."""
    from nect_amd.sampling import ct_sampling as cs

    angle, perm, jitter, offset, _ = w.batch(0)
    args = (perm, w.c_geometry, angle, w.S, w.B, 0, w.step_mm, True, 0.0, w.device)
    kw = {"jitter": jitter, "detector_offset": offset, "return_counts": True}
    pts, dist, cnt = cs.sample(*args, **kw)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        cs.sample(*args, **kw)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    bytes_ = 20 * w.B + 12 * w.S * w.B + 8 * w.B
    rate = bytes_ / (ms / 1e3) / 1e9
    return {"ms": round(ms, 4), "achieved": round(rate, 2), "unit": "GB/s", "frac": round(rate / HBM_PEAK_GBS, 4),
            "kept_samples": int(cnt.sum().item())}


def recompute_vs_store(dev, B: int = 1 << 18, S: int = 192, steps: int = 3):
    """Training step on the parallel-beam slab (512x512x64 voxels of 1 mm, detector 64 rows x 512, hash grid
    16x2 T=2^20, MLP 2x64, FP16), 2^18 rays of one view (pixels drawn with replacement) x <= 192 samples:
    the fused path, which recomputes each sample from a 64-byte ray descriptor, against the reference's
    structure run on this library's kernels -- points [B, S, 3] stored by the sampler, zero points masked
    and compacted, features stored for the backward by the autograd modules, torch Adam."""
    from nect_amd import tcnn
    from nect_amd.fused import FusedTrainStep
    from nect_amd.sampling import ct_sampling as cs
    from nect_amd.sampling.geometry import Geometry

    geo = Geometry(nDetector=(64, 512), dDetector=(1.0, 1.0), mode="parallel", nVoxel=(64, 512, 512),
                   dVoxel=(1.0, 1.0, 1.0), angles=[math.pi * i / 720 for i in range(720)])
    cg = geo.get_c_geometry()
    l_max = geo.max_distance_traveled
    step_mm = l_max / S
    enc = {"otype": "HashGrid", "n_levels": 16, "n_features_per_level": 2, "log2_hashmap_size": 20,
           "base_resolution": 16, "per_level_scale": 1.5}
    net = {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "Softplus", "n_neurons": 64,
           "n_hidden_layers": 2}
    from nect_amd import phantom

    model = tcnn.NetworkWithInputEncoding(3, 1, enc, net, dtype=torch.float16).to(dev)
    gen = torch.Generator(device=dev).manual_seed(5)
    slab = phantom.ellipse_slab(512, 64, 1.0, per_slice=200, seed=4)
    proj = phantom.project(slab, "parallel", (64, 512), (1.0, 1.0), 0.0, 0.0,
                           [geo.angles[k * 7] for k in range(steps + 1)], device=dev, reach=400.0)
    proj = (proj / proj.max()).reshape(steps + 1, -1)
    batches = []
    for k in range(steps + 1):
        idx = torch.randint(0, 64 * 512, (B,), device=dev, generator=gen)
        batches.append((geo.angles[k * 7], idx, torch.rand(B, device=dev, generator=gen) - 0.5,
                        torch.zeros((B, 2), device=dev), proj[k][idx].contiguous()))
    out = {"workload": "parallel slab 512x512x64 of random extruded ellipses (~200 per slice), 2^18 rays x "
                       "<=192 samples, hash grid 16x2 T=2^20, MLP 2x64, fp16"}

    # store: the reference's batch body (base_trainer.py:517-653) with this library's modules
    opt = torch.optim.Adam([model.params], lr=1e-3)

    def store_step(b):
        angle, idx, jit, off, tgt = b
        opt.zero_grad(set_to_none=True)
        pts, dist = cs.sample(idx, cg, angle, S, B, 0, step_mm, True, 0.0, dev, jitter=jit, detector_offset=off)
        flat = pts.view(-1, 3)
        mask = torch.all(flat == 0, dim=-1)
        mu = model(flat[~mask].contiguous())
        full = torch.zeros((B * S, 1), dtype=torch.float32, device=dev)
        full[~mask] = mu.float()
        y = full.view(B, S).sum(1) * (dist / l_max)
        loss = torch.nn.functional.mse_loss(y, tgt)
        loss.backward()
        opt.step()

    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats(dev)
    base = torch.cuda.memory_allocated(dev)
    store_step(batches[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        store_step(batches[k + 1])
    torch.cuda.synchronize()
    out["store_ms_per_step"] = round((time.perf_counter() - t0) / steps * 1e3, 2)
    out["store_peak_extra_gb"] = round((torch.cuda.max_memory_allocated(dev) - base) / 1e9, 3)
    params = model.params.detach().clone()
    del model, opt
    torch.cuda.empty_cache()

    # recompute: the fused path (FP16 as above, then BF16 and FP32 side by side, then FP16 under a 2 GB budget
    # for the per-sample buffers, where the batch runs in ray chunks)
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc

    grid = GridDesc(3, 16, 2, 20, 16, 1.5)
    mlp = MlpDesc(32, 1, 64, 2, "ReLU", "Softplus")

    def fused(dtype, budget=None):
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats(dev)
        base = torch.cuda.memory_allocated(dev)
        fs = FusedTrainStep(grid, mlp, params, dtype, max_rays=B, max_points_per_ray=S, loss="L2",
                            memory_budget=budget)
        angle, idx, jit, off, tgt = batches[0]
        fs.step(cg, angle, idx, 0, B, S, step_mm, True, jit, off, tgt, l_max, lr=1e-3)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            angle, idx, jit, off, tgt = batches[k + 1]
            fs.step(cg, angle, idx, 0, B, S, step_mm, True, jit, off, tgt, l_max, lr=1e-3)
        torch.cuda.synchronize()
        ms = round((time.perf_counter() - t0) / steps * 1e3, 2)
        extra = round((torch.cuda.max_memory_allocated(dev) - base) / 1e9, 3)
        res = (ms, extra, int(fs.last_samples()), fs.chunk_rays)
        del fs
        torch.cuda.empty_cache()
        return res

    ms, extra, samples, _ = fused(torch.float16)
    out["recompute_ms_per_step"], out["recompute_peak_extra_gb"], out["samples_per_step"] = ms, extra, samples
    out["speedup"] = round(out["store_ms_per_step"] / out["recompute_ms_per_step"], 2)
    out["recompute_bf16_ms_per_step"] = fused(torch.bfloat16)[0]
    out["recompute_fp32_ms_per_step"] = fused(torch.float32)[0]
    ms, extra, _, chunk = fused(torch.float16, 2 << 30)
    out["recompute_budget_2gb_ms_per_step"] = ms
    out["recompute_budget_2gb_peak_extra_gb"] = extra
    out["recompute_budget_2gb_chunk_rays"] = chunk
    return out


def train_4d_bf16(dev, B: int = 1 << 17, S: int = 128, steps: int = 10, train_steps: int = 200,
                  train_views: int = 32):
    """The fused BF16 step on the 4D workload (hash grid over x, y, z, t: 16x2, T=2^21; MLP 3x64): 256^3 cone
    volume (0.5 mm voxels), detector 512x512 at 0.5 mm, SOD 400, SDD 800, 2000 views over 5 turns, one t per
    view, 2^17 rays x <= 128 samples per step, targets from the brine-front sample's noisy projections
    (``Views4D``).  Timed: ``steps`` steps after 3 warm-up steps.  Then ``train_steps`` more steps over
    ``train_views`` views spread over the 2000 (untimed) train the model that ``serving_bench`` serves; the
    loss at the start and end of that training is reported."""
    from nect_amd.fused import FusedTrainStep, init_params
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc
    from nect_amd.sampling.geometry import Geometry

    n_views = 2000
    geo = Geometry(nDetector=(512, 512), dDetector=(0.5, 0.5), mode="cone", DSD=800.0, DSO=400.0,
                   nVoxel=(256, 256, 256), dVoxel=(0.5, 0.5, 0.5),
                   angles=[2 * math.pi * 5 * i / n_views for i in range(n_views)])
    cg = geo.get_c_geometry()
    l_max = geo.max_distance_traveled
    grid = GridDesc(4, 16, 2, 21, 16, 1.5)
    mlp = MlpDesc(32, 1, 64, 3, "ReLU", "Softplus")
    fs = FusedTrainStep(grid, mlp, init_params(grid, mlp, seed=3, device=dev), torch.bfloat16, max_rays=B,
                        max_points_per_ray=S, loss="L2")
    fs.profile = True
    gen = torch.Generator(device=dev).manual_seed(9)
    print("[bench] 4D sample projections", file=sys.stderr, flush=True)
    views = Views4D(dev, geo, n_views)
    timed_views = [(k * 37) % n_views for k in range(3 + steps)]
    for v in timed_views:
        views(v)

    def batch(v):
        idx = torch.randperm(512 * 512, device=dev, generator=gen)[:B]
        return (v, idx, torch.rand(B, device=dev, generator=gen) - 0.5, views(v)[idx].contiguous())

    def run(b, lr):
        v, idx, jit, tgt = b
        return fs.step(cg, geo.angles[v], idx, 0, B, S, l_max / S, True, jit, torch.zeros((B, 2), device=dev),
                       tgt, l_max, lr=lr, t=v / (n_views - 1))

    for k in range(3):
        run(batch(timed_views[k]), 1e-3)
    batches = [batch(v) for v in timed_views[3:]]
    torch.cuda.synchronize()
    fs.reset_timers()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for b in batches:
        run(b, 1e-3)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    out = {"workload": "4D cone, 256^3 brine-front sample (porosity ~0.35, 20% soluble grains, Poisson I0=5e3), "
                       "detector 512^2, 2^17 rays x <=128 samples, hash grid 4D 16x2 T=2^21, MLP 3x64, bf16",
           "ms_per_step": round(ms, 3), "rays_per_s": round(B / (ms / 1e3), 1),
           "samples_per_step": int(fs.last_samples()),
           "kernels": {k: round(v, 4) for k, v in fs.timer_means().items()}}
    if train_steps > 0:
        print(f"[bench] training the 4D model ({train_steps} steps)", file=sys.stderr, flush=True)
        fs.profile = False
        tv = [int(i * n_views / train_views) for i in range(train_views)]
        losses = []
        for k in range(train_steps):
            loss = run(batch(tv[(k * 7) % train_views]), 1e-2)
            if k < 10 or k >= train_steps - 10:
                losses.append(loss.clone())
        torch.cuda.synchronize()
        ls = [float(x) for x in losses]
        out["trained"] = {"steps": train_steps, "views": train_views, "lr": 1e-2,
                          "loss_first10": round(sum(ls[:10]) / 10, 6), "loss_last10": round(sum(ls[-10:]) / 10, 6)}
        _TRAINED_4D[str(dev)] = fs.params.detach().clone()
    return out


def static_preset_step(dev, w: "Workload", steps: int = 5, S: int = 384, points_per_batch: int = 5_000_000):
    """The reference's default static model as its preset configures it (cfg/static/hash_grid.yaml): hash grid 19
    levels x 4 features, T = 2^21, base 8, max resolution factor 2 (per-level scale (2*256/8)^(1/18) on the 256^3
    volume), network 6 x 64 LeakyReLU with ReLU output, L1 loss (cfg/default.yaml); the FP16 fused step on the
    256^3 cone workload's geometry and projections.  Per step the reference's Projector takes
    points_per_batch // 2 points (base_trainer.py:141-144; 5 M when sized automatically) as rays of
    points_per_ray samples (projector.py:92); S = 384 is where the preset's sample schedule ends on a 256^2
    detector (1.5 x 256)."""
    from nect_amd.fused import FusedTrainStep, init_params
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc

    grid = GridDesc(3, 19, 4, 21, 8, (2 * 256 / 8) ** (1 / 18))
    mlp = MlpDesc(19 * 4, 1, 64, 6, "LeakyReLU", "ReLU")
    B = (points_per_batch // 2) // S
    fs = FusedTrainStep(grid, mlp, init_params(grid, mlp, seed=1337, device=dev), torch.float16, max_rays=B,
                        max_points_per_ray=S, loss="L1")
    fs.profile = True
    gen = torch.Generator(device=dev).manual_seed(5)
    step_mm = w.l_max / S
    print("[bench] static preset", file=sys.stderr, flush=True)

    def batch(k):
        v = k % len(w.view_ids)
        perm = torch.randperm(w.B, device=dev, generator=gen)
        return (float(w.geometry.angles[w.view_ids[v]]), perm, torch.rand(B, device=dev, generator=gen) - 0.5,
                torch.zeros((B, 2), device=dev), w.proj[v][perm[:B]].contiguous())

    def run(b):
        angle, perm, jitter, offset, target = b
        fs.step(w.c_geometry, angle, perm, 0, B, S, step_mm, True, jitter, offset, target, w.l_max, lr=1e-3)

    for k in range(3):
        run(batch(k))
    batches = [batch(3 + k) for k in range(steps)]
    torch.cuda.synchronize()
    fs.reset_timers()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for b in batches:
        run(b)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    samples = int(fs.last_samples())
    return {"workload": "static preset (hash grid 19 levels x 4 features, T=2^21, base 8; MLP 6x64 LeakyReLU, ReLU "
                        "output; L1), fp16 fused step, 256^3 cone 256^2 detector, 2.5 M points per step (the "
                        "reference's Projector batch)",
            "rays_per_step": B, "samples_per_ray_max": S, "ms_per_step": round(ms, 2),
            "rays_per_s": round(B / (ms / 1e3), 1), "samples_per_s": round(samples / (ms / 1e3), 1),
            "samples_per_step": samples, "kernels": {k: round(v, 4) for k, v in fs.timer_means().items()}}


def quadcubes_step(dev, steps: int = 5, S: int = 256, points_per_batch: int = 5_000_000, split: bool = False):
    """The reference's default dynamic model as its preset configures it (cfg/dynamic/quadcubes.yaml): four 3-D
    hash grids over (z,y,x), (y,x,t), (x,z,t), (z,y,t) of 23 levels x 4 features, T = 2^23, base 16, max
    resolution factor 2, plus the (z,y,x,t) identity; network 4 x 128 LeakyReLU; FP16 mixed precision; the
    reference's per-step batch -- its Projector takes points_per_batch // 2 points (base_trainer.py:141-144;
    points_per_batch is 5 M when sized automatically) as rays of points_per_ray samples (projector.py:92), here
    <= 256 -- with the L1 loss of cfg/default.yaml, on the 4D workload's geometry and brine-front projections
    (Views4D).  Trained through ModuleTrainStep: patch-ordered
    batch, nested grids encoded in one launch, the layers as nect_gemm_act products with their activations and
    k_wgrad weight gradients, ray-group scatter per grid, dense Adam.  split: each grid's scatter as two launches
    (LDS-aggregated levels, then the direct levels), to time the phases apart."""
    from nect_amd.module_step import ModuleTrainStep
    from nect_amd.networks import QuadCubes
    from nect_amd.sampling.geometry import Geometry

    n_views = 2000
    geo = Geometry(nDetector=(512, 512), dDetector=(0.5, 0.5), mode="cone", DSD=800.0, DSO=400.0,
                   nVoxel=(256, 256, 256), dVoxel=(0.5, 0.5, 0.5),
                   angles=[2 * math.pi * 5 * i / n_views for i in range(n_views)])
    cg = geo.get_c_geometry()
    l_max = geo.max_distance_traveled
    enc = {"otype": "HashGrid", "n_levels": 23, "n_features_per_level": 4, "log2_hashmap_size": 23,
           "base_resolution": 16, "per_level_scale": (2 * 256 / 16) ** (1 / 22)}
    net = {"otype": "FullyFusedMLP", "activation": "LeakyReLU", "output_activation": "None", "n_neurons": 128,
           "n_hidden_layers": 4}
    model = QuadCubes(enc, net, include_identity=True, dtype=torch.float16).to(dev)
    B = (points_per_batch // 2) // S
    st = ModuleTrainStep(model, torch.float16, B, S, loss="L1", dynamic=True)
    st.scatter_split = split
    gen = torch.Generator(device=dev).manual_seed(21)
    print("[bench] QuadCubes preset", file=sys.stderr, flush=True)
    views = Views4D(dev, geo, n_views)
    vs = [(k * 53) % n_views for k in range(3 + steps)]

    def run(v):
        idx = torch.randperm(512 * 512, device=dev, generator=gen)[:B]
        return st.step(cg, geo.angles[v], idx, 0, B, S, l_max / S, True,
                       torch.rand(B, device=dev, generator=gen) - 0.5, torch.zeros((B, 2), device=dev),
                       views(v)[idx].contiguous(), l_max, lr=2e-4, t=v / (n_views - 1))

    for v in vs[:3]:
        run(v)
    for v in vs[3:]:
        views(v)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for v in vs[3:]:
        run(v)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    samples = st.last_samples()
    return {"workload": "QuadCubes preset (4 grids x 23 levels x 4 features, T=2^23, identity; MLP 4x128 "
                        "LeakyReLU, L1), fp16 mixed precision, 4D cone 512^2 detector, brine-front targets, "
                        "2.5 M points per step (the reference's Projector batch)",
            "rays_per_step": B, "samples_per_ray_max": S, "params": int(model.net.params.numel()),
            "ms_per_step": round(ms, 2), "rays_per_s": round(B / (ms / 1e3), 1),
            "samples_per_s": round(samples / (ms / 1e3), 1), "samples_per_step": samples,
            "amp": st.amp_status(), **({"scatter_split": True} if split else {})}


def psnr_bench(dev, steps: int, seed: int = 0):
    """Reconstruction quality of the FP16 fused path against the plain float32 torch path at equal steps.

    The 64^3 cone-beam case: 10 seeded ellipsoids (attenuation 0-1) in a 32 mm cube of 0.5 mm voxels,
    detector 64x64 at 1 mm, 90 views over 360 degrees, SOD 500 mm, SDD 1000 mm, analytic float64 line
    integrals (no noise) normalised by their maximum; hash grid 16x2 T=2^17, MLP 2x64 ReLU/softplus;
    4096 rays (one whole view) per step, <= 64 samples per ray, L2, Adam lr 1e-2.  Both paths start from
    the same parameters and see the same batches.  The reconstructions are evaluated at the voxel
    centres and compared with the voxelised phantom (PSNR over the phantom's range), and the final
    projections of four held-out angles with the analytic line integrals."""
    from nect_amd import phantom
    from nect_amd.baseline_torch import TorchTrainStep
    from nect_amd.fused import FusedTrainStep, init_params
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc
    from nect_amd.sampling.geometry import Geometry
    from nect_amd.serving import QueryRequest, QueryServer

    n_views, S, B = 90, 64, 64 * 64
    angles = [2 * math.pi * i / n_views for i in range(n_views)]
    geo = Geometry(nDetector=(64, 64), dDetector=(1.0, 1.0), mode="cone", DSD=1000.0, DSO=500.0,
                   nVoxel=(64, 64, 64), dVoxel=(0.5, 0.5, 0.5), angles=angles)
    cg = geo.get_c_geometry()
    l_max = geo.max_distance_traveled
    step_mm = l_max / S
    ph = phantom.ellipsoids((32.0, 32.0, 32.0), n=10, seed=seed)
    proj = phantom.project(ph, "cone", (64, 64), (1.0, 1.0), 500.0, 1000.0, angles, device=dev)
    pmax = float(proj.max())
    proj = (proj / pmax).reshape(n_views, -1).contiguous()
    truth = phantom.voxelize(ph, (64, 64, 64), (0.5, 0.5, 0.5), device=dev)

    grid = GridDesc(3, 16, 2, 17, 16, 1.5)
    mlp = MlpDesc(32, 1, 64, 2, "ReLU", "Softplus")
    p0 = init_params(grid, mlp, seed=1337, device=dev)
    fused = FusedTrainStep(grid, mlp, p0.clone(), torch.float16, max_rays=B, max_points_per_ray=S, loss="L2")
    plain = TorchTrainStep(grid, mlp, p0.clone(), lr=1e-2, loss="L2")
    gen = torch.Generator(device=dev).manual_seed(seed)
    order = torch.randint(0, n_views, (steps,), generator=torch.Generator().manual_seed(seed)).tolist()
    times = {"fp16_fused": 0.0, "fp32_torch": 0.0}
    print(f"[bench] psnr: {steps} steps of both paths", file=sys.stderr, flush=True)
    for k in range(steps):
        v = order[k]
        perm = torch.randperm(B, device=dev, generator=gen)
        jitter = torch.rand(B, device=dev, generator=gen) - 0.5
        offset = torch.zeros((B, 2), device=dev)
        target = proj[v][perm].contiguous()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fused.step(cg, angles[v], perm, 0, B, S, step_mm, True, jitter, offset, target, l_max, lr=1e-2)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        plain.step(cg, angles[v], perm, 0, B, S, step_mm, True, jitter, offset, target, l_max)
        torch.cuda.synchronize()
        times["fp16_fused"] += t1 - t0
        times["fp32_torch"] += time.perf_counter() - t1

    # reconstructions at the voxel centres, normalised coordinate (i + 0.5) / 64 on each axis
    c0, c1 = 0.5 / 64, 63.5 / 64
    req = QueryRequest((64, 64, 64), (c0, c0, c0), (c1, c1, c1), 0.0, l_max=l_max, span=pmax)
    sh = fused.shadow
    srv = QueryServer(grid, mlp, sh[: mlp.n_params], sh[mlp.n_params:], out_dtype=torch.float32)
    est16 = srv.run_alone(req).float()
    lin = torch.linspace(c0, c1, 64, device=dev)
    z, y, x = torch.meshgrid(lin, lin, lin, indexing="ij")
    with torch.no_grad():
        est32 = plain.model(torch.stack([z.flatten(), y.flatten(), x.flatten()], 1))[:, 0].view(64, 64, 64)
    est32 = (est32 / l_max) * pmax
    out = {"workload": "64^3 cone-beam, 10 ellipsoids, 90 views, 4096 rays x <=64 samples/step, L2, Adam 1e-2",
           "steps": steps,
           "psnr_db": {"fp16_fused": round(phantom.psnr(est16, truth), 3),
                       "fp32_torch": round(phantom.psnr(est32, truth), 3)},
           "train_s": {k: round(v, 2) for k, v in times.items()}}
    out["psnr_db"]["fp16_minus_fp32"] = round(out["psnr_db"]["fp16_fused"] - out["psnr_db"]["fp32_torch"], 3)

    # projections of four angles between the training views against the analytic integrals
    held = [2 * math.pi * (i + 0.5) / n_views for i in (3, 25, 47, 71)]
    ref = phantom.project(ph, "cone", (64, 64), (1.0, 1.0), 500.0, 1000.0, held, device=dev).reshape(4, -1) / pmax
    pix = torch.arange(B, device=dev)
    errs = {"fp16_fused": [], "fp32_torch": []}
    for j, a in enumerate(held):
        zeros = torch.zeros(B, device=dev)
        fused.forward_rays(cg, a, pix, 0, B, S, step_mm, True, zeros, torch.zeros((B, 2), device=dev))
        y16 = fused.predict(B, l_max)
        from nect_amd.baseline_torch import sample_points

        pts, dist = sample_points(cg, a, pix, 0, B, S, step_mm, True, zeros, torch.zeros((B, 2), device=dev))
        with torch.no_grad():
            mu = plain.model(pts.view(-1, 3)).view(B, S)
        mu = torch.where(torch.all(pts == 0, -1), torch.zeros((), device=dev), mu)
        y32 = mu.sum(1) * (dist / l_max)
        scale = ref[j].abs().max()
        errs["fp16_fused"].append(float((y16 - ref[j]).abs().mean() / scale))
        errs["fp32_torch"].append(float((y32 - ref[j]).abs().mean() / scale))
    out["heldout_projection_mean_abs_err_rel_to_max"] = {k: round(sum(v) / len(v), 5) for k, v in errs.items()}
    return out


def psnr_bench_256(dev, steps: int, seed: int = 0):
    """The FP16 fused path against the plain float32 torch path at equal steps on the 256^3 workload itself:
    the packed-sphere phantom (porosity ~0.35) seen through the 256^2 cone detector (SOD 400, SDD 800), Poisson
    noise at I0 = 1e4 then -log, normalised by the largest noisy value; hash grid 16x2 T=2^19, MLP 2x64
    ReLU/softplus, one whole view (2^16 rays x <= 256 samples) per step, distinct views in a seeded order, L2,
    Adam lr 1e-2.  Both paths start from the same parameters and see the same batches.  Reported: PSNR of the
    voxel-centre reconstructions against the voxelised phantom, and the error of three held-out projections
    against the analytic (noiseless) line integrals."""
    from nect_amd import phantom
    from nect_amd.baseline_torch import TorchTrainStep, sample_points
    from nect_amd.fused import FusedTrainStep, init_params
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc
    from nect_amd.sampling.geometry import Geometry
    from nect_amd.serving import QueryRequest, QueryServer

    n_views, S, B, N = 400, 256, 256 * 256, 256
    angles = [2 * math.pi * i / n_views for i in range(n_views)]
    geo = Geometry(nDetector=(256, 256), dDetector=(1.0, 1.0), mode="cone", DSD=800.0, DSO=400.0,
                   nVoxel=(N, N, N), dVoxel=(0.5, 0.5, 0.5), angles=angles)
    cg = geo.get_c_geometry()
    l_max = geo.max_distance_traveled
    step_mm = l_max / S
    ph, porosity = config2_phantom(dev)
    order = torch.randperm(n_views, generator=torch.Generator().manual_seed(seed)).tolist()
    train_views, held = order[:steps], order[-3:]
    print(f"[bench] psnr256: projecting {len(train_views)} views", file=sys.stderr, flush=True)
    clean = phantom.project(ph, "cone", (256, 256), (1.0, 1.0), 400.0, 800.0, [angles[v] for v in train_views],
                            device=dev)
    noisy = phantom.poisson_log(clean, I0_CONFIG2, generator=torch.Generator(device=dev).manual_seed(seed + 1))
    pnorm = float(noisy.max())
    targets = (noisy / pnorm).reshape(len(train_views), -1).contiguous()
    del clean, noisy
    truth = phantom.voxelize_spheres(ph, N, 0.5, device=dev)

    grid = GridDesc(3, 16, 2, 19, 16, 1.5)
    mlp = MlpDesc(32, 1, 64, 2, "ReLU", "Softplus")
    p0 = init_params(grid, mlp, seed=1337, device=dev)
    fused = FusedTrainStep(grid, mlp, p0.clone(), torch.float16, max_rays=B, max_points_per_ray=S, loss="L2")
    plain = TorchTrainStep(grid, mlp, p0.clone(), lr=1e-2, loss="L2")
    gen = torch.Generator(device=dev).manual_seed(seed)
    times = {"fp16_fused": 0.0, "fp32_torch": 0.0}
    print(f"[bench] psnr256: {steps} steps of both paths", file=sys.stderr, flush=True)
    for k, v in enumerate(train_views):
        perm = torch.randperm(B, device=dev, generator=gen)
        jitter = torch.rand(B, device=dev, generator=gen) - 0.5
        offset = torch.zeros((B, 2), device=dev)
        target = targets[k][perm].contiguous()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fused.step(cg, angles[v], perm, 0, B, S, step_mm, True, jitter, offset, target, l_max, lr=1e-2)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        plain.step(cg, angles[v], perm, 0, B, S, step_mm, True, jitter, offset, target, l_max)
        torch.cuda.synchronize()
        times["fp16_fused"] += t1 - t0
        times["fp32_torch"] += time.perf_counter() - t1
        if (k + 1) % 50 == 0:
            print(f"[bench] psnr256: step {k + 1}", file=sys.stderr, flush=True)

    # reconstructions at the voxel centres, in the phantom's attenuation units
    c0, c1 = 0.5 / N, 1 - 0.5 / N
    req = QueryRequest((N, N, N), (c0, c0, c0), (c1, c1, c1), 0.0, l_max=l_max, span=pnorm)
    sh = fused.shadow
    srv = QueryServer(grid, mlp, sh[: mlp.n_params], sh[mlp.n_params:], out_dtype=torch.float32)
    est16 = srv.run_alone(req).float()
    lin = torch.linspace(c0, c1, N, device=dev)
    yy, xx = torch.meshgrid(lin, lin, indexing="ij")
    est32 = torch.empty((N, N, N), device=dev)
    with torch.no_grad():
        for i in range(N):
            pts = torch.stack([torch.full_like(yy, float(lin[i])).flatten(), yy.flatten(), xx.flatten()], 1)
            est32[i] = plain.model(pts)[:, 0].view(N, N)
    est32 = (est32 / l_max) * pnorm
    out = {"workload": "256^3 cone-beam (detector 256^2, SOD 400, SDD 800), packed spheres porosity "
                       f"{porosity:.3f}, Poisson I0=1e4 then -log, 2^16 rays x <=256 samples/step, hash grid 16x2 "
                       "T=2^19, MLP 2x64, L2, Adam 1e-2, same parameters and batches",
           "steps": steps, "distinct_views": len(train_views),
           "psnr_db": {"fp16_fused": round(phantom.psnr(est16, truth), 3),
                       "fp32_torch": round(phantom.psnr(est32, truth), 3)},
           "train_s": {k: round(v, 2) for k, v in times.items()}}
    out["psnr_db"]["fp16_minus_fp32"] = round(out["psnr_db"]["fp16_fused"] - out["psnr_db"]["fp32_torch"], 3)
    del est16, est32, truth

    # three held-out views against the analytic noiseless line integrals (normalised like the targets)
    ref = phantom.project(ph, "cone", (256, 256), (1.0, 1.0), 400.0, 800.0, [angles[v] for v in held],
                          device=dev).reshape(len(held), -1) / pnorm
    pix = torch.arange(B, device=dev)
    errs = {"fp16_fused": [], "fp32_torch": []}
    for j, v in enumerate(held):
        zeros = torch.zeros(B, device=dev)
        fused.forward_rays(cg, angles[v], pix, 0, B, S, step_mm, True, zeros, torch.zeros((B, 2), device=dev))
        y16 = fused.predict(B, l_max)
        pts, dist = sample_points(cg, angles[v], pix, 0, B, S, step_mm, True, zeros, torch.zeros((B, 2), device=dev))
        with torch.no_grad():
            mu = torch.cat([plain.model(c.view(-1, 3)) for c in pts.split(8192)]).view(B, S)
        mu = torch.where(torch.all(pts == 0, -1), torch.zeros((), device=dev), mu)
        y32 = mu.sum(1) * (dist / l_max)
        scale = ref[j].abs().max()
        errs["fp16_fused"].append(float((y16 - ref[j]).abs().mean() / scale))
        errs["fp32_torch"].append(float((y32 - ref[j]).abs().mean() / scale))
    out["heldout_projection_mean_abs_err_rel_to_max"] = {k: round(sum(e) / len(e), 5) for k, e in errs.items()}
    return out


def serving_model(dev, timing: bool = False, host_copy: bool = False):
    """The serving model: 4D hash grid (x, y, z, t) 16x2, T=2^21, base 16, growth 1.5 + MLP 3x64 ReLU/Softplus,
    BF16 parameters -- those train_4d_bf16 trained on the brine-front sample when it ran in this process, else
    random -- FP16 output, in a QueryServer (2^16-voxel chunks, 32 per launch)."""
    from nect_amd.fused import init_params
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc
    from nect_amd.serving import QueryServer

    grid = GridDesc(4, 16, 2, 21, 16, 1.5)
    mlp = MlpDesc(32, 1, 64, 3, "ReLU", "Softplus")
    trained = _TRAINED_4D.get(str(dev))
    if trained is not None:
        params = trained
    else:
        params = init_params(grid, mlp, seed=11, device=dev)
        g = torch.Generator(device=dev).manual_seed(12)
        params[mlp.n_params:] = (torch.rand(grid.n_params, device=dev, generator=g) * 2 - 1) * 0.1
    shadow = params.to(torch.bfloat16)
    srv = QueryServer(grid, mlp, shadow[: mlp.n_params], shadow[mlp.n_params:], hbm_budget=64 << 30,
                      chunk=1 << 16, max_jobs=32, max_inflight=2, out_dtype=torch.float16, timing=timing,
                      host_copy=host_copy)
    srv.params = params
    srv.trained = trained is not None
    return srv


def query_volume(srv, reps: int = 3, traffic_file: str | None = None) -> dict:
    """A 512^3 volume served alone (after one warm-up), nect_query launches timed with HIP events on their
    stream; the roofline of k_query from the algorithmic work per voxel and the mean launch duration."""
    from nect_amd.serving import volume_request

    srv.run_alone(volume_request(0.5))
    srv.kernel_ms()
    walls = []
    for _ in range(reps):
        t0 = time.perf_counter()
        srv.run_alone(volume_request(0.5))
        walls.append((time.perf_counter() - t0) * 1e3)
    launches = srv.kernel_ms()
    kern_ms = sum(ms for ms, _ in launches) / reps
    vox = 512 ** 3
    gbs = QUERY_BYTES_PER_VOXEL * vox / (kern_ms / 1e3) / 1e9
    tflops = QUERY_FLOP_PER_VOXEL * vox / (kern_ms / 1e3) / 1e12
    measured = {}
    if traffic_file:
        try:
            with open(traffic_file) as f:
                measured = json.load(f)
        except (OSError, ValueError):
            measured = {}
    per_voxel = measured.get("query_bytes_per_voxel")
    return {"voxels": vox, "wall_ms": round(min(walls), 2), "kernel_ms": round(kern_ms, 2),
            "launches": len(launches) // reps, "mean_launch_ms": round(kern_ms * reps / len(launches), 4),
            "voxels_per_s": round(vox / (kern_ms / 1e3), 1),
            "roofline": {"kernel": "k_query", "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "traffic": round(per_voxel * vox / (len(launches) // reps)) if per_voxel else None,
                         "algorithmic_bytes_per_voxel": QUERY_BYTES_PER_VOXEL,
                         "mfma": {"achieved": round(tflops, 2), "peak": MFMA_PEAK_TFLOPS["bf16"], "unit": "TFLOP/s",
                                  "frac": round(tflops / MFMA_PEAK_TFLOPS["bf16"], 4),
                                  "flop_per_voxel": QUERY_FLOP_PER_VOXEL}}}


def serving_bench(dev, seconds: float, rates=(20, 50, 100), traffic_file: str | None = None):
    """Export queries against the serving model: seeded Poisson arrivals at each rate for ``seconds``, 70%
    512^2 slices / 25% sub-volumes of side 64..256 / 5% full 512^3 volumes, admitted by a 64 GB HBM budget;
    a request completes when its result is in pinned host memory (as the reference's exports end with .cpu());
    next to the sequential float32 torch export of the same request stream."""
    from nect_amd.baseline_torch import TorchExport
    from nect_amd.serving import mixed_workload, slice_request

    srv = serving_model(dev, timing=True, host_copy=True)
    print("[bench] serving warmup", file=sys.stderr, flush=True)
    srv.run_alone(slice_request("xy", 0.5, 0.5))
    alone = query_volume(srv, reps=2, traffic_file=traffic_file)
    srv.timing = False
    srv.serve(mixed_workload(20, 1.0, seed=99))   # fills the pinned-buffer pool
    model = ("trained on the 4D brine-front sample (train_4d_bf16)" if srv.trained else "random init")
    result = {"model": f"4D hash grid 16x2 T=2^21 + MLP 3x64, bf16 weights/table, fp16 output ({model})",
              "mix": "70% 512^2 slices, 25% sub-volumes side 64-256, 5% 512^3 volumes; Poisson arrivals",
              "latency": "arrival to result in pinned host memory (device-to-host copy included)",
              "arrival_window_s": seconds, "hbm_budget_gb": 64, "volume_alone": alone,
              "voxels_per_s_alone": alone["voxels_per_s"], "rates": {}}
    for rate in rates:
        print(f"[bench] serving {rate} req/s", file=sys.stderr, flush=True)
        stats = srv.serve(mixed_workload(rate, seconds, seed=rate))
        result["rates"][str(rate)] = {k: (round(x, 3) if isinstance(x, float) else x) for k, x in stats.items()}
    print("[bench] sequential fp32 torch export", file=sys.stderr, flush=True)
    base = TorchExport(srv.grid, srv.mlp, srv.params)
    base.run(slice_request("xy", 0.5, 0.5))
    torch.cuda.synchronize()
    bstats = base.serve(mixed_workload(rates[0], min(seconds, 2.0), seed=rates[0]))
    result["sequential_fp32_torch"] = {"rate": rates[0], "arrival_window_s": min(seconds, 2.0),
                                       **{k: (round(x, 3) if isinstance(x, float) else x) for k, x in bstats.items()}}
    return result


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


def time_cpu_baseline(w: Workload, grid, mlp, seconds: float):
    """The oracle's whole training step (C sampler, numpy hash-grid indices, torch autograd, torch.optim.Adam)
    in float32 on the host cores -- the PyTorch-CPU restatement of the reference path; bounded samples:
      * the 256^3 workload's step on a batch of 1024 rays x <= 256 samples, repeated for about `seconds`;
      * the 64^3 configuration's step (4096 rays x <= 64 samples, hash grid T=2^17), a few of its 2000 steps;
      * the same 256^3 sample in float64 (secondary) and the model on one 256^2 query slice."""
    import numpy as np

    from oracle import hashgrid as hg
    from oracle.geometry import OracleGeometry
    from oracle.train import CpuTrainer, model_forward
    from nect_amd.fused import init_params

    threads = min(16, os.cpu_count() or 1)  # the host-core share of one GPU
    torch.set_num_threads(threads)
    cfg = {"n_params": mlp.n_params, "n_input_padded": 32, "n_hidden_layers": 2, "n_output_padded": 16,
           "activation": "ReLU", "output_activation": "Softplus"}
    rng = np.random.default_rng(0)
    print("[bench] timing the CPU baseline", file=sys.stderr, flush=True)

    def timed(trainer, fields, n_pix, B, S, step_mm, l_max, angles, budget, min_steps=2):
        steps, t0 = 0, time.perf_counter()
        while steps < min_steps or time.perf_counter() - t0 < budget:
            idx = rng.permutation(n_pix).astype(np.int64)
            jit = (rng.random(B) - 0.5).astype(np.float32)
            trainer.step(fields, float(angles[steps % len(angles)]), idx, 0, B, S, step_mm, True, jit,
                         np.zeros((B, 2), np.float32), rng.random(B).astype(np.float32) * 0.3, l_max)
            steps += 1
        return steps, time.perf_counter() - t0

    # the 256^3 workload
    spec2 = hg.GridSpec(3, 16, 2, 19, 16, 1.5)
    fields2 = OracleGeometry(mode="cone", nDetector=(256, 256), dDetector=(1.0, 1.0), DSD=800.0, DSO=400.0,
                             nVoxel=(256, 256, 256), dVoxel=(0.5, 0.5, 0.5)).c_struct_fields()
    p2 = init_params(grid, mlp, seed=1337, device="cpu").double().numpy()
    B2 = 1024
    tr = CpuTrainer(spec2, cfg, p2, lr=1e-2, dtype=torch.float32)
    n2, dt2 = timed(tr, fields2, 256 * 256, B2, 256, w.step_mm, w.l_max, w.geometry.angles, seconds)
    tr64 = CpuTrainer(spec2, cfg, p2, lr=1e-2, dtype=torch.float64)
    n64, dt64 = timed(tr64, fields2, 256 * 256, B2, 256, w.step_mm, w.l_max, w.geometry.angles, 0.0)
    # the 64^3 configuration (its own grid)
    from nect_amd.grid import GridDesc

    g1 = GridDesc(3, 16, 2, 17, 16, 1.5)
    spec1 = hg.GridSpec(3, 16, 2, 17, 16, 1.5)
    og1 = OracleGeometry(mode="cone", nDetector=(64, 64), dDetector=(1.0, 1.0), DSD=1000.0, DSO=500.0,
                         nVoxel=(64, 64, 64), dVoxel=(0.5, 0.5, 0.5))
    p1 = init_params(g1, mlp, seed=1337, device="cpu").double().numpy()
    tr1 = CpuTrainer(spec1, cfg, p1, lr=1e-2, dtype=torch.float32)
    angles1 = [2 * math.pi * i / 90 for i in range(90)]
    n1, dt1 = timed(tr1, og1.c_struct_fields(), 64 * 64, 4096, 64, og1.max_distance_traveled / 64,
                    og1.max_distance_traveled, angles1, seconds / 2, min_steps=3)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next(line.split(":", 1)[1].strip() for line in f if line.startswith("model name"))
    except (OSError, StopIteration):
        pass
    n = 256
    lin = (np.arange(n, dtype=np.float32) + np.float32(0.5)) / np.float32(n)
    yy, xx = np.meshgrid(lin, lin, indexing="ij")
    pts = np.stack([np.full(n * n, np.float32(0.5)), yy.ravel(), xx.ravel()], 1).astype(np.float32)
    tq = time.perf_counter()
    with torch.no_grad():
        model_forward(spec2, cfg, tr.p.detach(), pts)
    query_s = time.perf_counter() - tq
    return {"value": round(n2 * B2 / dt2, 1), "unit": "rays/s", "cores": torch.get_num_threads(), "kind": "port",
            "dtype": "fp32", "os_cpu_count": os.cpu_count(), "cpu_model": model,
            "sample": f"{n2} full training steps (sampler, hash grid, MLP, integral, L2, backward, Adam) of the "
                      f"256^3 workload on {B2} rays x <=256 samples each, {dt2:.1f} s",
            "config1_64cube": {"rays_per_s": round(n1 * 4096 / dt1, 1), "s_per_step": round(dt1 / n1, 3),
                               "sample": f"{n1} of the 2000 training steps of the 64^3 configuration (4096 rays x "
                                         f"<=64 samples, hash grid T=2^17), {dt1:.1f} s; 2000 steps would take "
                                         f"{2000 * dt1 / n1 / 60:.0f} min"},
            "float64": {"rays_per_s": round(n64 * B2 / dt64, 1),
                        "sample": f"{n64} steps of the 256^3 sample in float64, {dt64:.1f} s"},
            "query_voxels_per_s": round(n * n / query_s, 1),
            "query_sample": f"the model on one {n}x{n} slice, {query_s:.2f} s"}


if __name__ == "__main__":
    main()
