"""snapshots_info.csv (no GPU): written in the reference driver's format and read back exactly."""
import ctypes

import numpy as np

from astaroth_amd import astaroth as ac


def test_snapshot_info_roundtrip(tmp_path):
    lib = ac.load("fp64")
    info = lib.config((24, 20, 16), AC_len=(2 * np.pi,) * 3)
    d = str(tmp_path).encode()
    t1 = 0.1 + 0.2  # not representable in %g: the exact column keeps every bit
    assert lib.acHostAppendSnapshotInfo(d, 0, 0, 0.0, ctypes.byref(info)) == 0
    assert lib.acHostAppendSnapshotInfo(d, 100, 1, t1, ctypes.byref(info)) == 0
    lines = (tmp_path / "snapshots_info.csv").read_text().splitlines()
    assert lines[0] == "use_double,mx,my,mz,step_number,modstep,t_step,t_step(exact)"
    assert lines[1].startswith("1,30,26,22,0,0,0,")
    assert lines[2].split(",")[:7] == ["1", "30", "26", "22", "100", "1", "0.3"]
    step, mod, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    assert lib.acHostReadSnapshotInfo(d, ctypes.byref(step), ctypes.byref(mod), ctypes.byref(t)) == 0
    assert (step.value, mod.value, t.value) == (100, 1, t1)
    # a new run (step 0) truncates the table
    assert lib.acHostAppendSnapshotInfo(d, 0, 0, 0.0, ctypes.byref(info)) == 0
    assert len((tmp_path / "snapshots_info.csv").read_text().splitlines()) == 2
    assert lib.acHostReadSnapshotInfo(str(tmp_path / "missing").encode(), None, None, None) == 1
