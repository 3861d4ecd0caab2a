"""The diagnostics timeseries format (no GPU): the reference's own timeseries file, written by
its standalone driver for test/subsonic_forced_turbulence, is read back column by column, and
rows written by acHostAppendTimeseries follow the same header and number format."""
import ctypes
import math
import os

import numpy as np
import pytest

from astaroth_amd import astaroth as ac
from astaroth_amd import timeseries as ts

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "subsonic_forced_turbulence_reference.ts")


def test_reads_reference_timeseries():
    names, data = ts.read_timeseries(GOLDEN)
    assert len(names) == 42 and data.shape == (26, 42)
    assert names[:3] == ["step", "t_step", "dt"]
    assert names[3:12] == ["uu_total_min", "uu_total_rms", "uu_total_max", "bb_total_min",
                           "bb_total_rms", "bb_total_max", "vA_total_min", "vA_total_rms",
                           "vA_total_max"]
    assert names[12:15] == ["VTXBUF_LNRHO_min", "VTXBUF_LNRHO_rms", "VTXBUF_LNRHO_max"]
    steps = data[:, 0]
    assert steps[0] == 0 and np.all(np.diff(steps[1:]) == 10)
    # min <= rms <= max for every magnitude group (|u|, |B|, vA)
    for c in (3, 6, 9):
        assert np.all(data[:, c] <= data[:, c + 1] + 1e-30)
        assert np.all(data[:, c + 1] <= data[:, c + 2] + 1e-30)


def diag_with(values):
    d = ac.AcDiagnostics()
    for k, v in values.items():
        if isinstance(v, (list, tuple)):
            getattr(d, k)[:] = v
        else:
            setattr(d, k, v)
    return d


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_writes_reference_format(prec, tmp_path):
    lib = ac.load(prec)
    path = tmp_path / "timeseries.ts"
    nf = ac.NUM_VTXBUF_HANDLES
    d = diag_with(dict(min=[-(q + 1) * 0.5 for q in range(nf)], max=[(q + 1) * 0.25 for q in range(nf)],
                       rms=[(q + 1) * 1e-3 for q in range(nf)], uu_min=1e-7, uu_rms=0.5,
                       uu_max=1.25, bb_min=0.0, bb_rms=3e-5, bb_max=4e-5, va_min=1e-9,
                       va_rms=2e-5, va_max=3.5e-5))
    assert not ts.append_row(lib, path, 0, 0.0, 0.4, d)
    assert not ts.append_row(lib, path, 10, 3.695833, 0.330443, d)
    text = path.read_text().splitlines()
    assert len(text) == 3
    ref_header = open(GOLDEN).readline().rstrip("\n")
    # same separators and leading columns as the reference's header line
    assert text[0].startswith(ref_header[:ref_header.index("VTXBUF_LNRHO_min")])
    names = [lib.lib.acGetVertexBufferName(q).decode() for q in range(nf)]
    assert names == ac.FIELD_NAMES
    assert text[0] == ref_header[:ref_header.index("VTXBUF_LNRHO_min")] + "".join(
        f"{n}_min  {n}_rms  {n}_max  " for n in names)
    assert text[2].startswith("10 3.695833e+00 3.304430e-01 1.000000e-07 5.000000e-01 1.250000e+00 ")
    cols, data = ts.read_timeseries(path)
    assert data.shape == (2, 12 + 3 * nf)
    assert data[1, cols.index("VTXBUF_AX_min")] == -2.5
    assert data[1, cols.index("vA_total_max")] == 3.5e-5
    # a NaN in a per-field reduction is reported (the reference driver stops on it)
    d.rms[3] = math.nan
    assert ts.append_row(lib, path, 20, 7.0, 0.3, d)
