"""Pins the MLP oracle: float64 autograd gradients agree with central finite differences."""
import torch

from oracle import mlp as om


def test_oracle_mlp_gradcheck():
    g = torch.Generator().manual_seed(0)
    for act, out_act in [("ReLU", "Softplus"), ("Tanh", "Sigmoid"), ("LeakyReLU", "None"), ("Squareplus", "Exponential")]:
        w = (torch.randn(16 * 8 + 8 * 8 + 16 * 8, generator=g, dtype=torch.float64) * 0.3).requires_grad_(True)
        x = torch.rand(5, 11, generator=g, dtype=torch.float64).requires_grad_(True)

        def f(w, x):
            # width 8 stand-in for the 64-wide production layer: same code path
            ws = [w[:128].view(8, 16), w[128:192].view(8, 8), w[192:].view(16, 8)]
            return om.forward(ws, x, act, out_act, 3)

        assert torch.autograd.gradcheck(f, (w, x), eps=1e-6, atol=1e-6)


def test_oracle_mlp_zero_padding_of_inputs():
    g = torch.Generator().manual_seed(1)
    ws = om.split(torch.randn(64 * 32 + 16 * 64, generator=g, dtype=torch.float64), 32, 64, 1, 16)
    x = torch.rand(4, 19, generator=g, dtype=torch.float64)
    y1 = om.forward(ws, x, "ReLU", "None", 1)
    y2 = om.forward(ws, torch.cat([x, torch.zeros(4, 13, dtype=torch.float64)], 1), "ReLU", "None", 1)
    assert torch.equal(y1, y2)
    # padded columns of W0 have no effect
    ws[0][:, 19:] = 123.0
    assert torch.equal(om.forward(ws, x, "ReLU", "None", 1), y1)
