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
    assert handle.nect_abi_version() == 1


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
