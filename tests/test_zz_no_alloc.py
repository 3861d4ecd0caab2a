"""acReduceNoAlloc (reductions with their partial results in a fixed device array, no allocation)
against the allocating entry points acDeviceReduceScal / Vec / VecScal and their
NoPostProcessing forms on the same device buffers: bit for bit, for every reduction type, in both
precisions, on two calls in a row whose second one runs on a smaller grid (the array keeps the
first call's partials). Also its argument checks."""
import ctypes
import sys

import numpy as np
import pytest

from astaroth_amd import astaroth as ac

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def noalloc(lib, name, ptrs, n, post, stream, out):
    arr = (ctypes.c_void_p * len(ptrs))(*ptrs)
    rc = lib.acReduceNoAlloc(ac.rtype(name), arr, len(ptrs), n[0], n[1], n[2], int(post),
                             ctypes.c_void_p(stream), ctypes.c_void_p(out.data_ptr()))
    return rc


def compare_on_grid(lib, n, seed):
    info = lib.config(n)
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(seed)
    mesh = ac.HostMesh(lib, info)
    assert lib.acHostMeshRandomize(ctypes.byref(mesh.struct)) == 0
    tdt = torch.float32 if lib.precision == "fp32" else torch.float64
    out = torch.zeros(1, dtype=tdt, device="cuda")
    with ac.Device(lib, info) as dev:
        dev.load_mesh(mesh)
        dev.periodic()
        dev.sync()
        stream = dev.stream_handle()
        ptr = [dev.field_ptrs(q)[0] for q in range(4)]
        cases = []
        for name in ("RTYPE_MAX", "RTYPE_MIN", "RTYPE_SUM", "RTYPE_RMS", "RTYPE_RMS_EXP"):
            for post in (True, False):
                cases.append((name, post, [ptr[0]], dev.reduce_scal(name, 0, post)))
                cases.append((name, post, ptr[:3], dev.reduce_vec(name, 0, 1, 2, post)))
        for name in ("RTYPE_ALFVEN_MAX", "RTYPE_ALFVEN_MIN", "RTYPE_ALFVEN_RMS"):
            for post in (True, False):
                cases.append((name, post, ptr[:4], dev.reduce_vec_scal(name, 0, 1, 2, 3, post)))
        for name, post, ptrs, want in cases:
            out.fill_(-1.0)
            torch.cuda.synchronize()
            assert noalloc(lib, name, ptrs, n, post, stream, out) == ac.AC_SUCCESS
            dev.sync()
            got = float(out.cpu()[0])
            say(f"  {lib.precision} {n} {name} post={post} nbufs={len(ptrs)}: {got!r} vs {want!r}")
            assert got == want, (name, post, len(ptrs), got, want)
        # argument checks: undefined arity, bad count, non-positive extent, null result
        arr = (ctypes.c_void_p * 4)(*ptr)
        assert lib.acReduceNoAlloc(ac.rtype("RTYPE_ALFVEN_MAX"), arr, 1, *n, 1,
                                   ctypes.c_void_p(stream),
                                   ctypes.c_void_p(out.data_ptr())) == ac.AC_FAILURE
        assert lib.acReduceNoAlloc(ac.rtype("RTYPE_MAX"), arr, 2, *n, 1, ctypes.c_void_p(stream),
                                   ctypes.c_void_p(out.data_ptr())) == ac.AC_FAILURE
        assert lib.acReduceNoAlloc(ac.rtype("RTYPE_MAX"), arr, 1, 0, n[1], n[2], 1,
                                   ctypes.c_void_p(stream),
                                   ctypes.c_void_p(out.data_ptr())) == ac.AC_FAILURE
        assert lib.acReduceNoAlloc(ac.rtype("RTYPE_MAX"), arr, 1, *n, 1, ctypes.c_void_p(stream),
                                   None) == ac.AC_FAILURE


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_noalloc_reductions_match_allocating_entry_points(prec, require_gpu):
    lib = ac.load(prec)
    say(f"case no-alloc reductions {prec}: 48^3 then 20 x 12 x 9")
    compare_on_grid(lib, (48, 48, 48), 11)
    compare_on_grid(lib, (20, 12, 9), 12)
