"""CPU check of the neighbour addressing of the z-marching adjoint transpose
(astaroth_amd/csrc/adjoint.hip, adjoint_transpose_z_kernel, its `ld`): a derivative-adjoint value
at offset (dx, dy) from column (i, j) in plane p of field k is read at

    g + sz (p + R) s + k count s + [colb + (unsigned)((dx + dy mx) s)]     (32-bit unsigned sum)

with colb = ((i + R) + mx (j + R)) s and s = sizeof(real). An earlier form added the signed
neighbour offset to the pointer as a zero-extended unsigned value, which moved negative offsets
4 GiB past the buffer. For every column, every offset the kernel uses (|dx|, |dy| <= 3 on the
axes and the diagonals), every plane of the march and every field, at grid sizes from 16^3 to
384^3 in both precisions, the 32-bit sum must not wrap and the address must be the intended
element inside the workspace of NG fields of m^3 values."""
import numpy as np
import pytest

R = 3
NG = 47  # derivative-adjoint fields in the workspace (mhd_adjoint.h)
OFFSETS = sorted({(d, 0) for d in range(-3, 4)} | {(0, d) for d in range(-3, 4)} |
                 {(d, d) for d in range(-3, 4)} | {(d, -d) for d in range(-3, 4)})


@pytest.mark.parametrize("size", [2, 4, 8])
@pytest.mark.parametrize("n", [(16, 16, 16), (64, 32, 48), (128, 128, 128), (256, 256, 256),
                               (384, 384, 384), (96, 8, 200)])
def test_transpose_neighbour_addresses(n, size):
    nx, ny, nz = n
    mx, my, mz = nx + 2 * R, ny + 2 * R, nz + 2 * R
    sz = mx * my
    count = sz * mz
    i, j = np.meshgrid(np.arange(nx, dtype=np.int64), np.arange(ny, dtype=np.int64), indexing="ij")
    colb = ((i + R) + mx * (j + R)) * size  # computed in int: must fit
    assert colb.max() < 2 ** 31
    workspace = NG * count * size
    for dx, dy in OFFSETS:
        rel = colb + (dx + dy * mx) * size          # exact value of the 32-bit unsigned sum
        off = (colb.astype(np.uint32) + np.uint32((dx + dy * mx) * size & 0xFFFFFFFF))
        assert rel.min() >= 0 and rel.max() < 2 ** 32, (n, dx, dy)
        assert np.array_equal(off.astype(np.int64), rel)
        for p in (-R, 0, nz // 2, nz - 1 + R):    # the first, interior and last planes read
            for k in (0, NG // 2, NG - 1):
                addr = sz * (p + R) * size + k * count * size + off.astype(np.int64)
                want = ((i + R + dx) + mx * (j + R + dy) + sz * (p + R) + k * count) * size
                assert np.array_equal(addr, want)
                assert addr.min() >= 0 and addr.max() + size <= workspace
