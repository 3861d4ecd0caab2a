"""Snapshots in the reference's on-disk format on the device: the files hold exactly the interior
of each field, and a run restarted from them continues bit-identically."""
import sys

import numpy as np
import pytest

from astaroth_amd import astaroth as ac

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_snapshot_write_read_restart(prec, tmp_path, require_gpu):
    print(f"case snapshot {prec}", flush=True)
    sys.stdout.flush()
    lib = ac.load(prec)
    n = (24, 20, 16)
    info = lib.config(n, AC_len=(2 * np.pi,) * 3, AC_cs_sound=1.0, AC_gamma=5.0 / 3.0,
                      AC_nu_visc=5e-3, AC_eta=5e-3, AC_chi=5e-3, AC_K_heatcond=5e-3,
                      AC_solve_tile_x=-1)  # one kernel on both devices: identical arithmetic
    dt = lib.np_real(2e-3)
    d = str(tmp_path).encode()
    with ac.Device(lib, info) as a, ac.Device(lib, info) as b:
        a.init_fourier(3, 16, 4, 0.05, 1e-4)
        for _ in range(3):
            a.step(dt)
        a.periodic()
        assert lib.acDeviceWriteSnapshot(a.handle, 0, d, b"3") == 0
        state = ac.HostMesh(lib, info)
        a.store_mesh(state)
        for q, name in enumerate(ac.FIELD_NAMES):
            raw = np.fromfile(tmp_path / f"{name}-3.mesh", dtype=lib.np_real)
            assert raw.size == n[0] * n[1] * n[2]
            want = state.data[q][3:-3, 3:-3, 3:-3].ravel()
            assert np.array_equal(raw, want), name
        assert lib.acDeviceReadSnapshot(b.handle, 0, d, b"3") == 0
        for dv in (a, b):
            for _ in range(4):
                dv.step(dt)
            dv.periodic()
        ra, rb = ac.HostMesh(lib, info), ac.HostMesh(lib, info)
        a.store_mesh(ra)
        b.store_mesh(rb)
        assert np.array_equal(ra.data, rb.data)
        # a file of the wrong size is an error, not a silent partial load
        (tmp_path / "VTXBUF_UUX-3.mesh").write_bytes(b"\0" * 16)
        assert lib.acDeviceReadSnapshot(b.handle, 0, d, b"3") == ac.AC_FAILURE
