"""CPU-side checks of the HIP library: it loads, exports every symbol include/astaroth.h
declares, and its host-side functions (parameters, config parser, host meshes, host periodic
fill, Fourier initial condition) behave as specified. No kernel is launched here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from astaroth_amd import astaroth as ac
import oracle.oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(REPO, "include", "astaroth.h")).read()
    body = text.split("AC_BEGIN_C_DECLARATIONS", 1)[1]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\b(ac[A-Z]\w*)\s*\(", body)
    return sorted(set(names))


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_library_exports_header(prec):
    lib = ac.load(prec)
    out = subprocess.run(["nm", "-D", "--defined-only", lib.path], capture_output=True,
                         text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [n for n in header_functions() if n not in exported]
    assert not missing, missing
    assert set(header_functions()) <= set(ac.EXPORTED)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_real_size_and_build(prec):
    lib = ac.load(prec)
    assert lib.lib.acGetRealSize() == (4 if prec == "fp32" else 8)
    assert b"gfx950" in lib.acGetBuildInfo()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_struct_layout(prec):
    lib = ac.load(prec)
    lib.lib.acGetMeshInfoSize.restype = ctypes.c_size_t
    lib.lib.acGetMeshSize.restype = ctypes.c_size_t
    assert ctypes.sizeof(lib.AcMeshInfo) == lib.lib.acGetMeshInfoSize()
    assert ctypes.sizeof(lib.AcMesh) == lib.lib.acGetMeshSize()


def test_default_config_and_derived_params():
    lib = ac.load("fp64")
    info = lib.config((24, 20, 16), AC_len=(2 * np.pi,) * 3)
    assert info["AC_nlocal"] == (24, 20, 16)
    assert info["AC_mlocal"] == (30, 26, 22)
    assert info["AC_nmin"] == (3, 3, 3)
    assert info["AC_nlocal_max"] == (27, 23, 19)
    np.testing.assert_allclose(info["AC_ds"], (2 * np.pi / 24, 2 * np.pi / 20, 2 * np.pi / 16))
    np.testing.assert_allclose(info["AC_inv_ds"], (24 / (2 * np.pi), 20 / (2 * np.pi),
                                                   16 / (2 * np.pi)))
    assert info["AC_cs2_sound"] == info["AC_cs_sound"] ** 2
    assert info["AC_mu0"] == pytest.approx(1.4)
    assert info["AC_lupwd"] and info["AC_lentropy"]


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_single_mesh_dims_setter(prec, tmp_path):
    """A caller that sets only one of the grid / local extents gets a device of that size: one
    process holds the whole grid (ref src/core/update_params/update_params.cc:173-177 pushes
    the grid dims into AC_nlocal in the build without MPI)."""
    lib = ac.load(prec)
    info = lib.AcMeshInfo()
    assert lib.acSetDefaultConfig(ctypes.byref(info)) == 0
    assert lib.acSetGridMeshDims(64, 48, 40, ctypes.byref(info)) == 0
    assert info["AC_ngrid"] == (64, 48, 40)
    assert info["AC_nlocal"] == (64, 48, 40)
    assert info["AC_mlocal"] == (70, 54, 46)
    assert info["AC_nlocal_max"] == (67, 51, 43)

    p = tmp_path / "grid.conf"
    p.write_text("AC_ngrid = {32, 32, 32}\nAC_len = {6.283185307179586, 6.283185307179586, "
                 "6.283185307179586}\n")
    info = lib.AcMeshInfo()
    assert lib.acLoadConfig(str(p).encode(), ctypes.byref(info)) == 0
    assert info["AC_nlocal"] == (32, 32, 32)
    assert lib.acSetLocalMeshDims(16, 24, 8, ctypes.byref(info)) == 0
    assert info["AC_nlocal"] == (16, 24, 8)
    assert info["AC_ngrid"] == (16, 24, 8)
    assert info["AC_mlocal"] == (22, 30, 14)
    # ds follows the extent actually held: len / n
    np.testing.assert_allclose(info["AC_ds"], (2 * np.pi / 16, 2 * np.pi / 24, 2 * np.pi / 8),
                               rtol=1e-6)


def test_config_loader(tmp_path):
    lib = ac.load("fp64")
    p = tmp_path / "test.conf"
    p.write_text("""/* block comment
    AC_nu_visc = 99 */
    AC_nu_visc = 5e-3  // trailing comment
    AC_eta = 4e-3
    AC_ngrid = {12, 14, 16}
    AC_len = {6.283185307179586, 6.283185307179586, 6.283185307179586}
    AC_lupwd = F
    AC_periodic_grid = {True,True,True}
    AC_unknown_param_of_other_subsystem = 3
    AC_kahan_bits = 8
    """)
    info = lib.AcMeshInfo()
    assert lib.acLoadConfig(str(p).encode(), ctypes.byref(info)) == 0
    assert info["AC_nu_visc"] == 5e-3
    assert info["AC_eta"] == 4e-3
    assert info["AC_nlocal"] == (12, 14, 16)
    assert not info["AC_lupwd"]
    assert info["AC_kahan_bits"] == 8
    assert info["AC_ds"][2] == pytest.approx(2 * np.pi / 16)
    assert lib.acLoadConfig(str(tmp_path / "missing.conf").encode(), ctypes.byref(info)) == 1


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_host_periodic_bounds_match_oracle(prec):
    lib = ac.load(prec)
    info = lib.config((7, 5, 6))
    mesh = ac.HostMesh(lib, info)
    rng = np.random.default_rng(1)
    mesh.data[...] = rng.random(mesh.data.shape)
    ref = mesh.data.copy()
    assert lib.acHostMeshApplyPeriodicBounds(ctypes.byref(mesh.struct)) == 0
    O.periodic_fill(ref, O.config_from_info(info))
    np.testing.assert_array_equal(mesh.data, ref)


def test_fourier_initial_condition():
    lib = ac.load("fp64")
    n = 16
    info = lib.config((n, n, n), AC_len=(2 * np.pi,) * 3)
    m1, m2 = ac.HostMesh(lib, info), ac.HostMesh(lib, info)
    assert lib.acHostMeshInitFourier(ctypes.byref(m1.struct), 1234, 16, 4, 0.05, 0.0) == 0
    assert lib.acHostMeshInitFourier(ctypes.byref(m2.struct), 1234, 16, 4, 0.05, 0.0) == 0
    np.testing.assert_array_equal(m1.data, m2.data)  # deterministic
    d = m1.data
    assert np.all(d[ac.VTXBUF_LNRHO] == 0) and np.all(d[ac.VTXBUF_ENTROPY] == 0)
    # each of uu, aa is a sum of 16 unit-polarised modes of amplitude 0.05
    for base in (ac.VTXBUF_UUX, ac.VTXBUF_AX):
        u = d[base:base + 3, 3:n + 3, 3:n + 3, 3:n + 3]
        assert 0 < np.max(np.abs(u)) <= 16 * 0.05 + 1e-12
        # spectral check: energy only at |k| <= 4, divergence-free (k . u_hat = 0)
        uh = np.fft.fftn(u, axes=(1, 2, 3))
        k = np.fft.fftfreq(n, 1.0 / n)
        KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
        k2 = KX ** 2 + KY ** 2 + KZ ** 2
        power = np.sum(np.abs(uh) ** 2, axis=0)
        assert np.all(power[k2 > 16] < 1e-18 * power.sum())
        div = KX * uh[0] + KY * uh[1] + KZ * uh[2]
        assert np.max(np.abs(div)) < 1e-9 * np.sqrt(power.max())
        # mean energy: each mode contributes 0.05^2/2 (unless two draws share a wavevector)
        e = np.mean(np.sum(u ** 2, axis=0))
        assert 0.2 * 16 * 0.05 ** 2 / 2 < e < 4 * 16 * 0.05 ** 2 / 2
    # halos are periodic images
    ref = d.copy()
    O.periodic_fill(ref, O.config_from_info(info))
    np.testing.assert_array_equal(d, ref)
    # lnrho perturbation is seeded, bounded and zero-mean-ish
    m3 = ac.HostMesh(lib, info)
    lib.acHostMeshInitFourier(ctypes.byref(m3.struct), 99, 4, 8, 0.1, 1e-4)
    l = m3.data[0, 3:n + 3, 3:n + 3, 3:n + 3]
    assert np.max(np.abs(l)) <= 1e-4 and np.std(l) > 3e-5


def test_no_gpu_create_fails_cleanly():
    """Without a visible GPU acDeviceCreate must fail with AC_FAILURE (no CPU fallback)."""
    from tests.conftest import gpu_available
    if gpu_available():
        pytest.skip("GPU present")
    lib = ac.load("fp32")
    info = lib.config((8, 8, 8))
    h = ctypes.c_void_p()
    assert lib.acDeviceCreate(0, info, ctypes.byref(h)) == ac.AC_FAILURE
    assert not h.value
