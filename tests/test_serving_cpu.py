"""Host logic of the export server: the brick traversal's codes and lengths, and that positions map onto
the grid one to one (the kernel's mapping, restated in numpy)."""
import numpy as np

from nect_amd.serving import traversal


def _voxels(n, code, length):
    """(iz, iy, ix, valid) for every traversal position, as the query kernel maps them."""
    v = np.arange(length, dtype=np.int64)
    if code == 0:
        iz, r = np.divmod(v, n[1] * n[2])
        iy, ix = np.divmod(r, n[2])
        return iz, iy, ix, iz < n[0]
    lz, ly, lx = (code >> 8) & 15, (code >> 4) & 15, code & 15
    ib = lz + ly + lx
    b, inner = v >> ib, v & ((1 << ib) - 1)
    nbx, nby = -(-n[2] // (1 << lx)), -(-n[1] // (1 << ly))
    bz, r = np.divmod(b, nby * nbx)
    by, bx = np.divmod(r, nbx)
    iz = (bz << lz) + (inner >> (ly + lx))
    iy = (by << ly) + ((inner >> lx) & ((1 << ly) - 1))
    ix = (bx << lx) + (inner & ((1 << lx) - 1))
    return iz, iy, ix, (iz < n[0]) & (iy < n[1]) & (ix < n[2])


def test_traversal_codes():
    assert traversal((512, 512, 512)) == (0x222, 512 ** 3)
    assert traversal((1, 512, 512)) == (0x022, 512 ** 2)
    assert traversal((512, 512, 1)) == (0x220, 512 ** 2)
    assert traversal((7, 33, 40)) == (0x222, 8 * 36 * 40)
    assert traversal((3, 2, 1)) == (0x110, 4 * 2 * 1)
    assert traversal((1, 1, 1)) == (0, 1)
    assert traversal((64, 64, 64), bricks=False) == (0, 64 ** 3)


def test_traversal_covers_every_voxel_once():
    for n in [(7, 33, 40), (1, 512, 9), (5, 1, 3), (65, 66, 67), (4, 4, 4), (2, 3, 17)]:
        code, length = traversal(n)
        iz, iy, ix, ok = _voxels(n, code, length)
        flat = (iz[ok] * n[1] + iy[ok]) * n[2] + ix[ok]
        assert np.array_equal(np.sort(flat), np.arange(n[0] * n[1] * n[2])), n
        # a position past the traversal maps outside the grid
        iz2, iy2, ix2, ok2 = _voxels(n, code, length + 64)
        assert not ok2[length:].any()
