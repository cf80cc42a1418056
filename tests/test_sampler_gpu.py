"""GPU parity of the HIP sampler against the oracle: bit-exact points, distances, counts."""
import numpy as np
import pytest
import torch

from oracle.geometry import OracleGeometry
from oracle.sampler import oracle_sample
from tests.helpers import cfg1_geometry, cfg2_geometry, cfg4_geometry, product_struct, seeded_batch

pytestmark = pytest.mark.gpu


def run_both(fields, angle, idx, B, start, S, step, uniform, jit, off, device):
    from nect_amd.sampling import ct_sampling as cs

    p_ref, d_ref, c_ref = oracle_sample(fields, angle, idx, B, start, S, step, uniform, jit, off)
    g = product_struct(fields)
    pts, dist, cnt = cs.sample(torch.from_numpy(idx).to(device), g, angle, S, B, start, step, uniform, 0.0, device,
                               jitter=torch.from_numpy(jit).to(device), detector_offset=torch.from_numpy(off).to(device),
                               return_counts=True)
    torch.cuda.synchronize()
    return (p_ref, d_ref, c_ref), (pts.cpu().numpy(), dist.cpu().numpy(), cnt.cpu().numpy())


def assert_same(ref, got, tag):
    p_ref, d_ref, c_ref = ref
    p, d, c = got
    assert np.array_equal(c, c_ref), f"{tag}: counts differ at {np.nonzero(c != c_ref)[0][:10]}"
    assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32)), f"{tag}: distances differ"
    bad = np.nonzero((p.view(np.uint32) != p_ref.view(np.uint32)).any(axis=(1, 2)))[0]
    assert bad.size == 0, f"{tag}: points differ on rays {bad[:10]}"


@pytest.mark.parametrize("angle", [0.0, 0.7853982, 2.0, 3.14159265, 5.5])
@pytest.mark.parametrize("offset", [0.0, 0.5])
def test_cfg1_cone_bitexact(device, angle, offset):
    g = cfg1_geometry()
    f = g.c_struct_fields()
    S = 64
    idx, jit, off = seeded_batch(64 * 64, 4096, seed=int(angle * 10) + int(offset * 2), det_offset_scale=offset)
    print(f"[sampler] cfg1 angle={angle} offset={offset}", flush=True)
    ref, got = run_both(f, angle, idx, 4096, 0, S, g.max_distance_traveled / S, True, jit, off, device)
    assert_same(ref, got, "cfg1")
    assert ref[2].max() > 0


def test_cfg2_full_batch_bitexact(device):
    g = cfg2_geometry()
    f = g.c_struct_fields()
    S = 256
    B = 65536
    idx, jit, off = seeded_batch(256 * 256, B, seed=5)
    print("[sampler] cfg2 B=65536 S=256", flush=True)
    ref, got = run_both(f, 1.234, idx, B, 0, S, g.max_distance_traveled / S, True, jit, off, device)
    assert_same(ref, got, "cfg2")


def test_cfg4_parallel_bitexact(device):
    g = cfg4_geometry()
    f = g.c_struct_fields()
    S = 192
    B = 8192
    idx, jit, off = seeded_batch(64 * 512, B, seed=9)
    print("[sampler] cfg4 parallel", flush=True)
    for angle in (0.0, 0.9, 1.5707964, 2.6):
        ref, got = run_both(f, angle, idx, B, 4096, S, g.max_distance_traveled / S, True, jit, off, device)
        assert_same(ref, got, f"cfg4 angle={angle}")


@pytest.mark.parametrize("uniform", [True, False])
def test_detector_rotation_cor_offset_nonuniform(device, uniform):
    g = OracleGeometry(mode="cone", nDetector=(48, 40), dDetector=(0.8, 0.9), DSD=700.0, DSO=300.0,
                       nVoxel=(40, 48, 44), dVoxel=(0.4, 0.45, 0.5), COR=0.37, offDetector=(1.1, -0.6),
                       rotDetector=(0.01, -0.02, 0.015))
    f = g.c_struct_fields()
    S = 80
    B = 48 * 40
    idx, jit, off = seeded_batch(B, B, seed=21, det_offset_scale=0.5)
    print(f"[sampler] rotated detector uniform={uniform}", flush=True)
    ref, got = run_both(f, 0.61, idx, B, 0, S, g.max_distance_traveled / S, uniform, jit, off, device)
    assert_same(ref, got, "rot")


@pytest.mark.parametrize("mode", ["cone", "parallel"])
def test_cylindrical_bitexact(device, mode):
    kw = dict(DSD=500.0, DSO=250.0) if mode == "cone" else {}
    g = OracleGeometry(mode=mode, nDetector=(32, 48), dDetector=(1.0, 1.0), nVoxel=(32, 48, 48),
                       dVoxel=(0.5, 0.5, 0.5), reconstruction_mode="cylindrical", **kw)
    f = g.c_struct_fields()
    S = 96
    B = 32 * 48
    idx, jit, off = seeded_batch(B, B, seed=33)
    print(f"[sampler] cylindrical {mode}", flush=True)
    ref, got = run_both(f, 1.1, idx, B, 0, S, g.max_distance_traveled / S, True, jit, off, device)
    assert_same(ref, got, "cyl")


def test_tail_and_out_of_range_slots(device):
    g = cfg1_geometry()
    f = g.c_struct_fields()
    idx = np.array([5, 10, 4096, 7000, 20, 4095, -3], dtype=np.int64)
    B = 10
    jit = np.linspace(-0.5, 0.49, B).astype(np.float32)
    off = np.zeros((B, 2), np.float32)
    print("[sampler] tail slots", flush=True)
    ref, got = run_both(f, 0.3, idx, B, 0, 64, 0.7, True, jit, off, device)
    assert_same(ref, got, "tail")
    ref, got = run_both(f, 0.3, idx, B, 3, 64, 0.7, True, jit, off, device)
    assert_same(ref, got, "tail-start3")


def test_fused_ray_descriptors_match_points(device):
    """Positions recomputed from 64-byte descriptors equal the materialised points."""
    from nect_amd.sampling import ct_sampling as cs

    g = cfg2_geometry()
    f = g.c_struct_fields()
    S = 256
    B = 16384
    idx, jit, off = seeded_batch(256 * 256, B, seed=44)
    step = g.max_distance_traveled / S
    print("[sampler] fused descriptors", flush=True)
    p_ref, d_ref, c_ref = oracle_sample(f, 2.2, idx, B, 0, S, step, True, jit, off)
    gs = product_struct(f)
    t_idx = torch.from_numpy(idx).to(device)
    rays = cs.ray_setup(t_idx, gs, 2.2, S, B, 0, step, True, torch.from_numpy(jit).to(device),
                        torch.from_numpy(off).to(device))
    max_tiles = cs.tiles_needed(B, S)
    tile_off, tile_ray, n_tiles = cs.ray_tiles(rays, max_tiles)
    pos, valid = cs.ray_positions(rays, tile_off, tile_ray, n_tiles, max_tiles)
    torch.cuda.synchronize()
    _, _, cnt, dist = cs.ray_fields(rays)
    assert np.array_equal(cnt.cpu().numpy(), c_ref)
    assert np.array_equal(dist.cpu().numpy(), d_ref)
    to = tile_off.cpu().numpy()
    tiles_per = (c_ref + 31) // 32
    assert np.array_equal(np.diff(to), tiles_per) and to[0] == 0
    assert int(n_tiles.item()) == int(tiles_per.sum())
    pos = pos.cpu().numpy()
    valid = valid.cpu().numpy()
    # slot order -> [ray, i]
    for r in np.random.default_rng(0).choice(B, 512, replace=False):
        n = c_ref[r]
        s0 = to[r] * 32
        got = pos[s0:s0 + n]
        assert np.array_equal(got.view(np.uint32), p_ref[r, :n].view(np.uint32)), r
        keep = ~np.all(p_ref[r, :n] == 0, axis=1)
        assert np.array_equal(valid[s0:s0 + n].astype(bool), keep)
        assert not valid[s0 + n:to[r + 1] * 32].any()
