"""GPU parity of the hash-grid encoder: indices/weights bit-exact, features and gradients within tolerance."""
import numpy as np
import pytest
import torch

from oracle import hashgrid as hg

pytestmark = pytest.mark.gpu

SPECS = {
    "cfg1": (3, 16, 2, 17, 16, 1.5),
    "cfg2": (3, 16, 2, 19, 16, 1.5),
    "cfg3_4d": (4, 16, 2, 21, 16, 1.5),
    "default_yaml": (3, 19, 4, 21, 8, 1.3459),
}


def points(n, D, seed, edge=False):
    rng = np.random.default_rng(seed)
    x = rng.random((n, D), dtype=np.float32)
    if edge:
        x[: n // 8] = np.float32(0.0)
        x[n // 8: n // 4] = np.float32(1.0)
        x[n // 4: n // 3, 0] = np.float32(0.5)
    return x


@pytest.mark.parametrize("name", list(SPECS))
def test_indices_and_weights_bitexact(device, name):
    from nect_amd.grid import GridDesc, index

    args = SPECS[name]
    print(f"[grid] index {name}", flush=True)
    d = GridDesc(*args)
    s = hg.GridSpec(*args)
    from tests.helpers import rounding_edge_points

    # random points (where a two-rounding position would give other weights on ~0.4% of the coordinates) and
    # points next to every level's cell boundaries
    x = np.concatenate([points(3000, args[0], seed=1, edge=True), rounding_edge_points(s, 32, seed=2)])
    idx, w = index(d, torch.from_numpy(x).to(device))
    torch.cuda.synchronize()
    i_ref, w_ref = hg.corners(s, x)
    assert np.array_equal(idx.cpu().numpy().view(np.uint32), i_ref)
    assert np.array_equal(w.cpu().numpy().view(np.uint32), w_ref.view(np.uint32))


@pytest.mark.parametrize("name", ["cfg1", "cfg3_4d", "default_yaml"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float16, torch.bfloat16])
def test_encode_matches_oracle(device, name, dtype):
    from nect_amd.grid import GridDesc, encode

    args = SPECS[name]
    print(f"[grid] encode {name} {dtype}", flush=True)
    d = GridDesc(*args)
    s = hg.GridSpec(*args)
    rng = np.random.default_rng(2)
    table32 = (rng.standard_normal(s.n_params) * 0.1).astype(np.float32)
    table = torch.from_numpy(table32).to(device).to(dtype)
    x = points(4099, args[0], seed=3, edge=True)
    out = encode(d, table, torch.from_numpy(x).to(device))
    torch.cuda.synchronize()
    ref = hg.encode(s, table.float().cpu().numpy(), x)
    got = out.float().cpu().numpy()
    # float accumulation of 2^D products, then rounding of the output to dtype
    tol = {torch.float32: 1e-5, torch.float16: 1e-3, torch.bfloat16: 8e-3}[dtype]
    assert np.allclose(got, ref, rtol=tol, atol=tol * 0.1), np.abs(got - ref).max()


@pytest.mark.parametrize("name", ["cfg1", "cfg3_4d", "default_yaml"])
def test_backward_matches_oracle(device, name):
    from nect_amd.grid import GridDesc, backward

    args = SPECS[name]
    print(f"[grid] backward {name}", flush=True)
    d = GridDesc(*args)
    s = hg.GridSpec(*args)
    x = points(2048, args[0], seed=4, edge=True)
    rng = np.random.default_rng(5)
    dout = rng.standard_normal((2048, d.n_output_dims)).astype(np.float32)
    grad = torch.zeros(d.n_params, dtype=torch.float32, device=device)
    backward(d, torch.from_numpy(x).to(device), torch.from_numpy(dout).to(device), grad)
    torch.cuda.synchronize()
    ref = hg.backward(s, x, dout)
    got = grad.cpu().numpy()
    scale = np.abs(ref).max()
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-6 * scale), np.abs(got - ref).max()
    assert np.array_equal(got != 0, ref != 0) or np.abs(got[(got != 0) != (ref != 0)]).max() < 1e-6 * scale


def test_strided_inputs_and_outputs(device):
    """Composite encodings read a column slice of the input and write a column slice of the output."""
    from nect_amd.grid import GridDesc, encode

    args = SPECS["cfg1"]
    d = GridDesc(*args)
    s = hg.GridSpec(*args)
    rng = np.random.default_rng(6)
    table = torch.from_numpy((rng.standard_normal(s.n_params) * 0.1).astype(np.float32)).to(device)
    big = torch.from_numpy(rng.random((777, 7), dtype=np.float32)).to(device)
    out = torch.zeros((777, 40), dtype=torch.float32, device=device)
    encode(d, table, big[:, 2:5], out=out[:, 4:36])
    torch.cuda.synchronize()
    ref = hg.encode(s, table.cpu().numpy(), big[:, 2:5].cpu().numpy())
    got = out.cpu().numpy()
    assert np.allclose(got[:, 4:36], ref, rtol=1e-5, atol=1e-7)
    assert np.all(got[:, :4] == 0) and np.all(got[:, 36:] == 0)


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.float32])
def test_grouped_feature_layout_is_a_permutation_of_rows(device, dtype):
    """nect_grid_encode_rays with grouped=1 writes the same bits as the row layout, plane g holding the
    features of levels g, 15-g (16 levels) slot by slot."""
    from nect_amd import _lib
    from nect_amd.fused import FusedTrainStep, init_params
    from nect_amd.grid import GridDesc
    from nect_amd.mlp import MlpDesc
    from tests.helpers import cfg1_geometry, product_struct, seeded_batch

    print(f"[grid] grouped layout {dtype}", flush=True)
    g = cfg1_geometry()
    f = g.c_struct_fields()
    S, B = 64, 512
    grid = GridDesc(3, 16, 2, 17, 16, 1.5)
    mlp = MlpDesc(32, 1, 64, 2, "ReLU", "Softplus")
    params = init_params(grid, mlp, seed=0, device=device)
    params[mlp.n_params:] = (torch.rand(grid.n_params, generator=torch.Generator().manual_seed(1)) - 0.5).to(device)
    fs = FusedTrainStep(grid, mlp, params, dtype, max_rays=B, max_points_per_ray=S)
    idx, jit, off = seeded_batch(64 * 64, B, seed=2)
    max_tiles = fs.forward_rays(product_struct(f), 0.4, torch.from_numpy(idx).to(device), 0, B, S,
                                g.max_distance_traveled / S, True, torch.from_numpy(jit).to(device),
                                torch.from_numpy(off).to(device))
    slots = max_tiles * 32
    st = _lib.stream_of(fs.device)
    rows = torch.full((slots, 32), 7.0, dtype=dtype, device=device)
    planes = torch.full((8, slots, 4), 7.0, dtype=dtype, device=device)
    for out, grouped in ((rows, 0), (planes, 1)):
        _lib.call("nect_grid_encode_rays", grid.struct_ref(), _lib.dtype_code(dtype), _lib.ptr(fs.table),
                  _lib.ptr(fs.rays[:B]), _lib.ptr(fs.tile_offsets), _lib.ptr(fs.tile_ray), _lib.ptr(fs.n_tiles),
                  int(max_tiles), 0.0, _lib.ptr(out), 32, grouped, st)
    torch.cuda.synchronize()
    used = int(fs.n_tiles.item()) * 32
    for grp in range(8):
        for m, level in enumerate((grp, 15 - grp)):
            a = rows[:used, 2 * level: 2 * level + 2]
            b = planes[grp, :used, 2 * m: 2 * m + 2]
            assert torch.equal(a.view(torch.int16 if dtype != torch.float32 else torch.int32),
                               b.view(torch.int16 if dtype != torch.float32 else torch.int32)), (grp, level)
