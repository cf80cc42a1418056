"""CPU checks of the C-ABI library: it loads and exports every symbol include/*.h declares."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(nect_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "nect_sample" in names and "nect_last_error" in names
    assert len(names) >= 6


def test_library_exports_every_declared_symbol():
    import torch  # noqa: F401  (binds the HIP runtime first)

    from nect_amd import _lib

    handle = _lib.lib()
    for name in declared_symbols():
        assert hasattr(handle, name), name
    assert set(_lib.exported_symbols()) <= set(declared_symbols())
    version = re.search(r"#define NECT_ABI_VERSION (\d+)", (ROOT / "include" / "nect_amd.h").read_text())
    assert handle.nect_abi_version() == int(version.group(1))


def test_error_path_without_gpu():
    """Host-side validation fails before any launch, with a readable message."""
    import torch  # noqa: F401

    from nect_amd import _lib

    g = _lib.NectGeometry()
    g.kind = 9
    g.n_det[0] = g.n_det[1] = 4
    rc = _lib.lib().nect_sample(ctypes.byref(g), 0.0, None, 0, 1, 0, 8, 1.0, 1, None, None, None, None, None, None)
    assert rc != 0
    assert b"geometry kind" in _lib.lib().nect_last_error()
    g.kind = 0
    rc = _lib.lib().nect_sample(ctypes.byref(g), 0.0, None, 0, 1, 0, 0, 1.0, 1, None, None, None, None, None, None)
    assert rc != 0 and b"num_points_per_ray" in _lib.lib().nect_last_error()


def test_grid_level_table_matches_oracle():
    """nect_grid_init (host-only) builds the same level table as the oracle's independent restatement."""
    import numpy as np

    from nect_amd.grid import GridDesc
    from oracle.hashgrid import GridSpec

    for args in [(3, 16, 2, 17, 16, 1.5), (3, 16, 2, 19, 16, 1.5), (4, 16, 2, 21, 16, 1.5), (3, 16, 2, 20, 16, 1.5),
                 (3, 19, 4, 21, 8, 1.3459), (4, 7, 4, 15, 3, 2.1)]:
        d = GridDesc(*args)
        s = GridSpec(*args)
        assert np.array_equal(np.array(d.scale, np.float32), s.scale), args
        assert d.resolution == list(s.resolution), args
        assert d.level_size == list(s.level_size), args
        assert d.offset == list(s.offset), args
        assert d.n_params == s.n_params, args


def test_no_kernel_takes_a_reference_parameter():
    """A kernel parameter of reference type would make the device dereference a host address."""
    import shutil
    import subprocess

    from nect_amd import _lib

    if shutil.which("c++filt") is None:
        import pytest

        pytest.skip("c++filt not available")
    data = _lib.LIB_PATH.read_bytes()
    names = sorted({m.decode() for m in re.findall(rb"_ZN4nect\d+k_[A-Za-z0-9_]+", data)})
    assert len(names) > 20
    demangled = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True).stdout
    kernels = [d for d in demangled.splitlines() if "::k_" in d]
    assert kernels and not [k for k in kernels if "&" in k], [k for k in kernels if "&" in k][:3]


def test_row_pitch_widens_only_whole_even_line_rows():
    """Feature buffers written a few bytes per row at a time get 16 more bytes per row exactly when a row would
    be a whole even number of 128-byte lines; other widths keep their rows."""
    import torch

    from nect_amd.tcnn import _row_pitch

    assert _row_pitch(384, torch.float16) == 392     # 768 B = six lines
    assert _row_pitch(128, torch.float16) == 136     # 256 B = two lines
    assert _row_pitch(376, torch.float16) == 376     # 752 B
    assert _row_pitch(384, torch.float32) == 388     # 1536 B = twelve lines
    assert _row_pitch(372, torch.float32) == 372
    assert _row_pitch(32, torch.float16) == 32       # 64 B
    for w, dt in ((384, torch.float16), (128, torch.float16), (384, torch.float32)):
        assert (_row_pitch(w, dt) * torch.empty((), dtype=dt).element_size()) % 16 == 0


def test_network_kind_follows_what_the_fused_kernels_cover():
    """64-neuron networks take the fused kernels (16-bit: up to 8 hidden layers and 128 padded inputs; float32:
    up to 3 layers and 64 inputs, the float32 backward's limit); anything else the GEMM network."""
    import torch

    from nect_amd.mlp import GemmMlpDesc, MlpDesc
    from nect_amd.tcnn import _mlp_desc

    cfg = lambda h, w=64: {"otype": "FullyFusedMLP", "n_neurons": w, "n_hidden_layers": h}  # noqa: E731
    assert isinstance(_mlp_desc(32, 1, cfg(6), torch.float16), MlpDesc)
    assert isinstance(_mlp_desc(32, 1, cfg(6), torch.bfloat16), MlpDesc)
    assert isinstance(_mlp_desc(32, 1, cfg(3), torch.float32), MlpDesc)
    assert isinstance(_mlp_desc(32, 1, cfg(4), torch.float32), GemmMlpDesc)
    assert isinstance(_mlp_desc(76, 1, cfg(2), torch.float32), GemmMlpDesc)
    assert isinstance(_mlp_desc(76, 1, cfg(2), torch.float16), MlpDesc)
    assert isinstance(_mlp_desc(372, 1, cfg(4, 128), torch.float16), GemmMlpDesc)
