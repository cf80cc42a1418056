"""GPU parity of the fused MFMA MLP against the float64 oracle (forward, weight and input gradients)."""
import numpy as np
import pytest
import torch

from oracle import mlp as om

pytestmark = pytest.mark.gpu

TOL = {torch.float32: 2e-5, torch.float16: 6e-3, torch.bfloat16: 4e-2}


def make(desc, dtype, seed, n, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    w = torch.randn(desc.n_params, generator=g) * (scale / np.sqrt(64))
    x = torch.rand(n, desc.n_input, generator=g) * 2 - 1
    # the oracle sees exactly the rounded values the kernel sees
    w = w.to(dtype).double()
    x = x.to(dtype).double()
    return w, x


def rel_err(got, ref):
    return (got - ref).abs().max().item() / max(ref.abs().max().item(), 1e-30)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float16, torch.bfloat16])
@pytest.mark.parametrize("hidden,act,out_act,n_in,n", [
    (2, "ReLU", "Softplus", 32, 4096),
    (3, "ReLU", "None", 32, 1000),
    (1, "LeakyReLU", "Sigmoid", 16, 77),
    (2, "Tanh", "ReLU", 19, 333),
    (2, "Softplus", "Exponential", 48, 129),
    (4, "ReLU", "None", 32, 700),
    (6, "LeakyReLU", "ReLU", 32, 1500),
    (6, "LeakyReLU", "ReLU", 76, 900),
])
def test_forward(device, dtype, hidden, act, out_act, n_in, n):
    from nect_amd.mlp import MlpDesc, forward

    print(f"[mlp] forward {dtype} H={hidden} {act}/{out_act} n_in={n_in} n={n}", flush=True)
    d = MlpDesc(n_in, 1, 64, hidden, act, out_act)
    w, x = make(d, dtype, seed=hidden * 7 + n_in, n=n)
    y = forward(d, w.to(dtype).to(device), x.to(dtype).to(device).contiguous())
    torch.cuda.synchronize()
    ws = om.split(w, d.n_input_padded, 64, hidden, d.n_output_padded)
    ref = om.forward(ws, x, act, out_act, 1, round_dtype=None if dtype == torch.float32 else dtype)
    assert y.shape == (n, 1)
    assert rel_err(y.cpu().double(), ref) < TOL[dtype], rel_err(y.cpu().double(), ref)


def test_forward_multiple_outputs_and_stride(device):
    from nect_amd.mlp import MlpDesc, forward

    d = MlpDesc(32, 5, 64, 2, "ReLU", "None")
    w, x = make(d, torch.float16, seed=3, n=500)
    big = torch.zeros((500, 40), dtype=torch.float16)
    big[:, :32] = x.half()
    out = torch.zeros((500, 8), dtype=torch.float32, device=device)
    forward(d, w.half().to(device), big.to(device)[:, :32], out=out[:, 1:6])
    torch.cuda.synchronize()
    ref = om.forward(om.split(w, 32, 64, 2, 16), x, "ReLU", "None", 5)
    got = out.cpu().double()
    assert rel_err(got[:, 1:6], ref) < TOL[torch.float16]
    assert torch.all(got[:, 0] == 0) and torch.all(got[:, 6:] == 0)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float16, torch.bfloat16])
@pytest.mark.parametrize("hidden,act,out_act,n_in,n", [
    (2, "ReLU", "Softplus", 32, 3000),
    (3, "ReLU", "None", 32, 700),
    (1, "LeakyReLU", "Sigmoid", 16, 77),
    (2, "Tanh", "Softplus", 48, 200),
    (4, "ReLU", "None", 32, 700),
    (6, "LeakyReLU", "ReLU", 32, 1500),
    (5, "Tanh", "Softplus", 64, 300),
    (6, "LeakyReLU", "ReLU", 76, 900),
    (2, "ReLU", "Softplus", 128, 500),
])
def test_backward(device, dtype, hidden, act, out_act, n_in, n):
    from nect_amd.mlp import MlpDesc, backward

    print(f"[mlp] backward {dtype} H={hidden} {act}/{out_act} n_in={n_in} n={n}", flush=True)
    d = MlpDesc(n_in, 1, 64, hidden, act, out_act)
    if dtype == torch.float32 and (hidden > 3 or n_in > 64):
        w, x = make(d, dtype, seed=1, n=n)
        with pytest.raises(RuntimeError, match="at most 3 hidden layers|wider than 64"):
            backward(d, w.float().to(device), x.float().to(device).contiguous(), torch.zeros(n, 1, device=device),
                     torch.zeros(d.n_params, device=device))
        return
    w, x = make(d, dtype, seed=hidden * 11 + n_in, n=n)
    g = torch.Generator().manual_seed(5)
    dy = torch.randn(n, 1, generator=g).double() / n
    # oracle gradients
    wr = w.clone().requires_grad_(True)
    xr = x.clone().requires_grad_(True)
    ref_y = om.forward(om.split(wr, d.n_input_padded, 64, hidden, d.n_output_padded), xr, act, out_act, 1,
                       round_dtype=None if dtype == torch.float32 else dtype)
    (ref_y * dy).sum().backward()
    dw = torch.zeros(d.n_params, dtype=torch.float32, device=device)
    dx = torch.zeros((n, n_in), dtype=torch.float32, device=device)
    loss_scale = 1.0 if dtype == torch.float32 else 1024.0 * n
    backward(d, w.to(dtype).to(device), x.to(dtype).to(device).contiguous(), dy.float().to(device), dw, dx,
             loss_scale=loss_scale, grad_scale=1.0)
    torch.cuda.synchronize()
    # remaining differences: 16-bit rounding of the back-propagated values fed to the matrix cores
    tol = {torch.float32: 2e-5, torch.float16: 6e-3, torch.bfloat16: 5e-2}[dtype]
    e_w = rel_err(dw.cpu().double(), wr.grad)
    e_x = rel_err(dx.cpu().double(), xr.grad)
    assert e_w < tol, e_w
    assert e_x < tol, e_x


def test_backward_accumulates_and_scales(device):
    from nect_amd.mlp import MlpDesc, backward

    d = MlpDesc(32, 1, 64, 2, "ReLU", "None")
    w, x = make(d, torch.float32, seed=9, n=256)
    dy = torch.ones(256, 1, dtype=torch.float32)
    dw1 = torch.zeros(d.n_params, device=device)
    backward(d, w.float().to(device), x.float().to(device), dy.to(device), dw1)
    dw2 = torch.full((d.n_params,), 1.0, device=device)
    backward(d, w.float().to(device), x.float().to(device), dy.to(device), dw2, grad_scale=0.5)
    torch.cuda.synchronize()
    assert torch.allclose(dw2.cpu(), 1.0 + 0.5 * dw1.cpu(), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_wgrad_matches_float64(device, dtype):
    """nect_wgrad (weight gradient over the samples, split across workgroups, float atomics) equals the float64
    product of the same 16-bit operands: ragged sample counts and output shapes (edges of the 128 x 128 tiles
    and of the 32-sample steps), accumulation into dw, the scale."""
    from nect_amd import _lib

    print(f"[mlp] wgrad {dtype}", flush=True)
    g = torch.Generator(device=device).manual_seed(5)
    for K, m, n in [(100_003, 128, 384), (4097, 16, 128), (70_000, 128, 128), (33, 5, 7), (300, 130, 129)]:
        dz = (torch.randn(K, m, device=device, generator=g) * 0.5).to(dtype)
        h = (torch.randn(K, n, device=device, generator=g) * 0.5).to(dtype)
        ref = dz.double().t() @ h.double()
        dw = torch.full((m, n), 0.25, dtype=torch.float32, device=device)
        _lib.call("nect_wgrad", _lib.dtype_code(dtype), dz.data_ptr(), m, h.data_ptr(), n, K, m, n, 2.0,
                  dw.data_ptr(), n, _lib.stream_of(device))
        torch.cuda.synchronize()
        want = 0.25 + 2.0 * ref
        err = float((dw.double() - want).abs().max() / ref.abs().max())
        assert err < 2e-5, (K, m, n, err)


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_gemm_act_matches_float64(device, dtype):
    """nect_gemm_act (a wide layer's product with its activation or act'(z) mask in the epilogue, rounded once)
    against the float64 product of the same 16-bit operands, rounded to the dtype: row counts that are not a
    multiple of the 64-row tile, K = 16 / 128 / 384 (A with a padded row pitch), M = 128 / 384, epilogues none,
    ReLU / LeakyReLU forward, and the masks from saved activations."""
    from nect_amd.mlp import gemm_act

    print(f"[mlp] gemm_act {dtype}", flush=True)
    g = torch.Generator(device=device).manual_seed(11)
    for n, k, m, pitch in [(100_003, 384, 128, 392), (4097, 128, 128, 128), (70, 16, 128, 16), (65, 128, 384, 136)]:
        a = (torch.randn(n, pitch, device=device, generator=g) * 0.5).to(dtype)[:, :k]
        bt = (torch.randn(m, k, device=device, generator=g) * 0.2).to(dtype)
        z = a.double() @ bt.double().t()
        hsaved = torch.randn(n, m, device=device, generator=g).to(dtype)
        for epi, slope in [(0, 1.0), (1, 0.0), (1, 0.01), (2, 0.0), (2, 0.01)]:
            got = gemm_act(a, bt, epi, slope, hsaved if epi == 2 else None)
            torch.cuda.synchronize()
            if epi == 0:
                want = z
            elif epi == 1:
                want = torch.where(z > 0, z, z * slope)
            else:
                want = torch.where(hsaved.double() > 0, z, z * slope)
            # one rounding of a float32 sum: half an ulp of the dtype (relative) plus the float32 accumulation
            # error (absolute, ~1e-6 here: measured against outputs of typical size 2, floored at 0.05)
            ulp = 2.0 ** -10 if dtype == torch.float16 else 2.0 ** -7
            err = (got.double() - want).abs() / (want.abs() + 0.05)
            assert float(err.max()) < 1.5 * ulp, (n, k, m, epi, slope, float(err.max()))
            if epi == 2 and slope == 0.0:
                assert torch.all(got[hsaved <= 0] == 0)
