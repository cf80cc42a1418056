"""Pins the oracle's composite models (oracle/networks.py) on the CPU."""
import numpy as np
import torch

from oracle import hashgrid as hg
from oracle import networks as on
from oracle.train import model_forward


def _cfg(n_in, width, hidden, act="LeakyReLU", out_act="None"):
    P = (n_in + 15) // 16 * 16
    n = width * P + (hidden - 1) * width * width + 16 * width
    return {"n_params": n, "n_input_padded": P, "width": width, "n_hidden_layers": hidden, "n_output_padded": 16,
            "activation": act, "output_activation": out_act}


def test_one_nested_grid_is_the_plain_grid_model():
    spec = hg.GridSpec(3, 8, 2, 12, 8, 1.4)
    cfg = _cfg(16, 64, 2, "ReLU", "Softplus")
    rng = np.random.default_rng(0)
    p = torch.from_numpy(rng.normal(0, 0.3, cfg["n_params"] + spec.n_params))
    x = rng.random((500, 3), dtype=np.float32)
    a = on.composite_forward([("grid", spec)], cfg, p, x)
    b = model_forward(spec, cfg, p, x)
    assert torch.equal(a, b)


def test_identity_columns_pass_through():
    cfg = _cfg(4, 16, 1, "None", "None")
    w = torch.zeros(cfg["n_params"], dtype=torch.float64)
    P = cfg["n_input_padded"]
    w[0 * P + 2] = 1.0                  # hidden 0 <- input column 2
    w[16 * P + 0] = 1.0                 # output <- hidden 0
    x = np.random.default_rng(1).random((50, 4), dtype=np.float32)
    y = on.composite_forward([("identity", 4)], cfg, w, x)
    assert torch.equal(y, torch.from_numpy(x[:, 2].astype(np.float64)))


def test_input_assembly_rows():
    zyx = np.array([[1.0, 2.0, 3.0]], np.float32)
    q = on.quad_inputs(zyx, 0.5)
    assert q.tolist() == [[1, 2, 3, 2, 3, 0.5, 3, 1, 0.5, 1, 2, 0.5, 1, 2, 3, 0.5]]
    assert on.quad_inputs(zyx, 0.5, identity=False).shape == (1, 12)
    h = on.hyper_inputs(zyx, 0.25)
    assert h.tolist() == [[1, 2, 3, 2, 3, 0.25, 3, 1, 0.25, 1, 2, 0.25, 1, 2, 3, 0.25, 1, 2, 3, 0.25]]


def test_composite_gradient_against_central_differences():
    s3, s4 = hg.GridSpec(3, 4, 2, 10, 4, 1.5), hg.GridSpec(4, 3, 2, 10, 4, 1.5)
    parts = [("grid", s3), ("grid", s3), ("grid", s4), ("identity", 4)]
    n_in = 3 * 2 + 4 + 4
    cfg = _cfg(4 * 2 * 2 + 3 * 2 + 4, 32, 2, "Tanh", "Softplus")
    rng = np.random.default_rng(2)
    n_par = cfg["n_params"] + 2 * s3.n_params + s4.n_params
    p = torch.from_numpy(rng.normal(0, 0.5, n_par)).requires_grad_(True)
    x = rng.random((40, n_in), dtype=np.float32)
    y = on.composite_forward(parts, cfg, p, x)
    (y * torch.linspace(-1, 1, 40, dtype=torch.float64)).sum().backward()
    g = p.grad.numpy()
    pick = np.concatenate([rng.choice(cfg["n_params"], 30, replace=False),
                           cfg["n_params"] + np.flatnonzero(g[cfg["n_params"]:])[:30]])
    h = 1e-6
    for i in pick:
        with torch.no_grad():
            q = p.detach().clone()
            q[i] += h
            fp = (on.composite_forward(parts, cfg, q, x) * torch.linspace(-1, 1, 40, dtype=torch.float64)).sum()
            q[i] -= 2 * h
            fm = (on.composite_forward(parts, cfg, q, x) * torch.linspace(-1, 1, 40, dtype=torch.float64)).sum()
        fd = float(fp - fm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
