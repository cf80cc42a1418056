"""nect_mlp_backward_noalloc -- the MLP backward with its per-workgroup partial weight gradients in the
library's own device array instead of a caller workspace -- against nect_mlp_backward, through the C ABI, on
two calls in a row (the second with fewer rows): bit-equal where it runs the same number of workgroups,
within float summation order where its scratch holds fewer."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _backward(entry, desc, dtype, w, x, dy, with_workspace):
    from nect_amd import _lib

    n = x.shape[0]
    dw = torch.zeros(desc.n_params, dtype=torch.float32, device=x.device)
    dx = torch.full((n, desc.n_input), float("nan"), dtype=torch.float32, device=x.device)
    args = [ctypes.byref(desc.struct), _lib.dtype_code(dtype), w.data_ptr(), x.data_ptr(), n, x.shape[1], None, None,
            dy.data_ptr(), 1, 64.0, None, 0.5, dw.data_ptr(), dx.data_ptr(), desc.n_input]
    if with_workspace:
        need = int(_lib.lib().nect_mlp_backward_workspace(ctypes.byref(desc.struct), _lib.dtype_code(dtype), n))
        ws = torch.empty(max(need, 16), dtype=torch.uint8, device=x.device)
        args += [ws.data_ptr(), ws.numel()]
    rc = getattr(_lib.lib(), entry)(*args, _lib.stream_of(x.device))
    assert rc == 0, _lib.lib().nect_last_error()
    torch.cuda.synchronize()
    return dw, dx


@pytest.mark.parametrize("hidden,n_in,dtype,same_blocks", [(2, 32, torch.float16, True),
                                                            (3, 32, torch.bfloat16, True),
                                                            (6, 76, torch.float16, False)])
def test_noalloc_backward_matches_workspace_backward(device, hidden, n_in, dtype, same_blocks):
    from nect_amd import _lib
    from nect_amd.mlp import MlpDesc

    print(f"[noalloc] mlp backward H={hidden} in={n_in} {dtype}", flush=True)
    _lib.lib().nect_mlp_backward_noalloc.restype = ctypes.c_int32
    desc = MlpDesc(n_in, 1, 64, hidden, "ReLU", "Softplus")
    g = torch.Generator().manual_seed(hidden)
    w = (torch.randn(desc.n_params, generator=g) * 0.2).to(device=device, dtype=dtype)
    for n in (200_003, 70_001):  # the second call is smaller and reuses the scratch the first one filled
        x = (torch.rand((n, n_in), generator=g) * 2 - 1).to(device=device, dtype=dtype)
        dy = torch.randn((n, 1), generator=g).to(device)
        dw_ref, dx_ref = _backward("nect_mlp_backward", desc, dtype, w, x, dy, True)
        dw, dx = _backward("nect_mlp_backward_noalloc", desc, dtype, w, x, dy, False)
        assert torch.isfinite(dw_ref).all() and torch.isfinite(dx_ref).all()
        assert torch.equal(dx, dx_ref)
        if same_blocks:
            assert torch.equal(dw, dw_ref), n
        else:
            err = float((dw - dw_ref).abs().max() / dw_ref.abs().max())
            assert err < 1e-5, (n, err)
