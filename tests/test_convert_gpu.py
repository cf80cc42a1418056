"""float32 -> bfloat16 / float16 conversion on the device (every table, activation and shadow copy the
kernels write in 16 bits goes through it): round to nearest, ties to even, subnormals kept, overflow to
infinity, NaN stays NaN.  Checked through `nect_cast` against known answers and against numpy on random bit
patterns of every exponent."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

# (float32 bits, bfloat16 bits) worked out by hand from the definition
BF16_KNOWN = [
    (0x3F800000, 0x3F80),  # 1.0
    (0x3F808000, 0x3F80),  # 1 + 2^-8: halfway, 0x3f80 is even
    (0x3F818000, 0x3F82),  # halfway above an odd mantissa: up
    (0x3F808001, 0x3F81),  # just above halfway
    (0x3F807FFF, 0x3F80),  # just below halfway
    (0xBF818000, 0xBF82),  # sign is kept
    (0x7F7FFFFF, 0x7F80),  # the largest float rounds to infinity
    (0x7F7F7FFF, 0x7F7F),  # the largest bfloat16 stays
    (0x00000001, 0x0000),  # the smallest subnormal rounds to zero
    (0x00008000, 0x0000),  # halfway between 0 and the smallest bfloat16 subnormal
    (0x00018000, 0x0002),  # halfway above an odd subnormal
    (0x00010000, 0x0001),  # the smallest bfloat16 subnormal
    (0x80000000, 0x8000),  # -0
    (0x7F800000, 0x7F80),  # +inf
    (0xFF800000, 0xFF80),  # -inf
]


def _bf16_rne(u: np.ndarray) -> np.ndarray:
    u = u.astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _patterns(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    u = rng.integers(0, 2**32, size=1 << 20, dtype=np.uint64).astype(np.uint32)
    # ties and near-ties on every exponent, subnormals included
    low = np.array([0x8000, 0x7FFF, 0x8001, 0x0000, 0xFFFF], dtype=np.uint32)
    u[: 5 * 4096] = (u[: 5 * 4096] & 0xFFFF0000) | np.repeat(low, 4096)
    u[5 * 4096: 6 * 4096] &= 0x807FFFFF  # subnormals
    return u


def _cast(device, u: np.ndarray, dtype: torch.dtype) -> np.ndarray:
    from nect_amd import _lib
    src = torch.from_numpy(u.view(np.float32).copy()).to(device)
    dst = torch.empty(src.numel(), dtype=dtype, device=device)
    _lib.call("nect_cast", src.data_ptr(), dst.data_ptr(), _lib.dtype_code(dtype), src.numel(),
              _lib.stream_of(device))
    torch.cuda.synchronize(device)
    return dst.view(torch.int16).cpu().numpy().view(np.uint16)


def test_bf16_known_answers(device):
    u = np.array([a for a, _ in BF16_KNOWN], dtype=np.uint32)
    got = _cast(device, u, torch.bfloat16)
    want = np.array([b for _, b in BF16_KNOWN], dtype=np.uint16)
    bad = [(hex(a), hex(g), hex(w)) for a, g, w in zip(u, got, want) if g != w]
    assert not bad, bad


def test_bf16_round_to_nearest_even(device):
    u = _patterns()
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    got = _cast(device, u, torch.bfloat16)
    assert np.array_equal(got[~nan], _bf16_rne(u[~nan])), int((got[~nan] != _bf16_rne(u[~nan])).sum())
    g = got[nan]
    assert np.all(((g & 0x7F80) == 0x7F80) & ((g & 0x7F) != 0)), "NaN inputs must stay NaN"


def test_f16_matches_numpy(device):
    u = _patterns(1)
    x = u.view(np.float32)
    nan = np.isnan(x)
    got = _cast(device, u, torch.float16)
    want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got[~nan], want[~nan]), int((got[~nan] != want[~nan]).sum())
    assert np.all(np.isnan(got[nan].view(np.float16)))
