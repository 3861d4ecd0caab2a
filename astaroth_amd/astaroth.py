"""ctypes mirror of the C API in include/astaroth.h.

`load("fp32")` / `load("fp64")` returns an `AstarothLib` whose attributes are the exported C
functions under their reference names (acDeviceCreate, acDeviceLoadMesh, ...), with argument
types set, plus the structures (AcMeshInfo, AcMesh, Volume, ...) sized for that build's AcReal.
`HostMesh` couples an AcMesh to a numpy array of shape (NUM_VTXBUF_HANDLES, mz, my, mx).

The library is the in-tree HIP build (astaroth_amd/lib); there is no CPU fallback: loading
raises when the shared library is missing, and device calls fail when no GPU is present.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

# ---- enums (same values as include/astaroth.h) ----
(VTXBUF_LNRHO, VTXBUF_UUX, VTXBUF_UUY, VTXBUF_UUZ, VTXBUF_AX, VTXBUF_AY, VTXBUF_AZ,
 VTXBUF_ENTROPY, NUM_VTXBUF_HANDLES) = range(9)
AC_TIMESTEP_COURANT, AC_TIMESTEP_REFERENCE = 0, 1

FIELD_NAMES = ["VTXBUF_LNRHO", "VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ", "VTXBUF_AX",
               "VTXBUF_AY", "VTXBUF_AZ", "VTXBUF_ENTROPY"]

INT_PARAMS = ["AC_max_steps", "AC_save_steps", "AC_start_step", "AC_kahan_bits",
              "AC_solve_tile_x", "AC_solve_tile_y", "AC_solve_zchunk", "AC_forcing_period_steps",
              "AC_solve_kernel"]
INT3_PARAMS = ["AC_ngrid", "AC_nlocal", "AC_mlocal", "AC_mgrid", "AC_nmin", "AC_nlocal_max",
               "AC_ngrid_max", "AC_multigpu_offset"]
REAL_PARAMS = ["AC_cdt", "AC_cdtv", "AC_cdts", "AC_nu_visc", "AC_cs_sound", "AC_zeta", "AC_eta",
               "AC_mu0", "AC_chi", "AC_cp_sound", "AC_gamma", "AC_lnT0", "AC_lnrho0",
               "AC_K_heatcond", "AC_cs2_sound", "AC_dsmin", "AC_current_time", "AC_dt",
               "AC_forcing_magnitude", "AC_relhel", "AC_kmin", "AC_kmax", "AC_forcing_phase",
               "AC_k_forcex", "AC_k_forcey", "AC_k_forcez", "AC_ff_hel_rex", "AC_ff_hel_rey",
               "AC_ff_hel_rez", "AC_ff_hel_imx", "AC_ff_hel_imy", "AC_ff_hel_imz", "AC_kaver",
               "AC_switch_forcing", "AC_forcing_period_t"]
REAL3_PARAMS = ["AC_ds", "AC_inv_ds", "AC_inv_ds_2", "AC_len"]
BOOL_PARAMS = ["AC_lupwd", "AC_lentropy", "AC_additive_timestep", "AC_periodic_grid_all",
               "AC_lforcing"]

for _names in (INT_PARAMS, INT3_PARAMS, REAL_PARAMS, REAL3_PARAMS, BOOL_PARAMS):
    for _i, _n in enumerate(_names):
        globals()[_n] = _i

AC_SUCCESS, AC_FAILURE, AC_NOT_ALLOCATED = 0, 1, 2
STREAM_DEFAULT = 0
STREAM_ALL = 32
NUM_STREAMS = 32

AC_REDUCE_OP_NO_REDUCE, AC_REDUCE_OP_MIN, AC_REDUCE_OP_MAX, AC_REDUCE_OP_SUM = range(4)
AC_NO_REDUCE_POST_PROCESSING, AC_RMS = 0, 1
(AC_NULL_KERNEL, AC_MAP_VTXBUF, AC_MAP_VTXBUF_SQUARE, AC_MAP_VTXBUF_EXP_SQUARE,
 AC_MAP_VTXBUF3_NORM, AC_MAP_VTXBUF3_SQUARE, AC_MAP_VTXBUF3_EXP_SQUARE,
 AC_MAP_VTXBUF4_ALFVEN_NORM, AC_MAP_VTXBUF4_ALFVEN_SQUARE) = range(9)

# (reduce_op, post, map_single, map_vec, map_vec_scal, name), include/astaroth.h RTYPE_*
RTYPES = {
    "RTYPE_MAX": (AC_REDUCE_OP_MAX, 0, AC_MAP_VTXBUF, AC_MAP_VTXBUF3_NORM, AC_NULL_KERNEL),
    "RTYPE_MIN": (AC_REDUCE_OP_MIN, 0, AC_MAP_VTXBUF, AC_MAP_VTXBUF3_NORM, AC_NULL_KERNEL),
    "RTYPE_SUM": (AC_REDUCE_OP_SUM, 0, AC_MAP_VTXBUF, AC_MAP_VTXBUF3_NORM, AC_NULL_KERNEL),
    "RTYPE_RMS": (AC_REDUCE_OP_SUM, 1, AC_MAP_VTXBUF_SQUARE, AC_MAP_VTXBUF3_SQUARE, AC_NULL_KERNEL),
    "RTYPE_RMS_EXP": (AC_REDUCE_OP_SUM, 1, AC_MAP_VTXBUF_EXP_SQUARE, AC_MAP_VTXBUF3_EXP_SQUARE,
                      AC_NULL_KERNEL),
    "RTYPE_ALFVEN_MAX": (AC_REDUCE_OP_MAX, 0, AC_NULL_KERNEL, AC_NULL_KERNEL,
                         AC_MAP_VTXBUF4_ALFVEN_NORM),
    "RTYPE_ALFVEN_MIN": (AC_REDUCE_OP_MIN, 0, AC_NULL_KERNEL, AC_NULL_KERNEL,
                         AC_MAP_VTXBUF4_ALFVEN_NORM),
    "RTYPE_ALFVEN_RMS": (AC_REDUCE_OP_SUM, 1, AC_NULL_KERNEL, AC_NULL_KERNEL,
                         AC_MAP_VTXBUF4_ALFVEN_SQUARE),
}

# Every symbol include/astaroth.h declares (checked by the CPU tests).
EXPORTED = [
    "acDeviceCreate", "acDeviceDestroy", "acDeviceResetMesh", "acDevicePrintInfo",
    "acDeviceSynchronizeStream", "acDeviceSwapBuffer", "acDeviceSwapBuffers",
    "acDeviceLoadScalarUniform", "acDeviceLoadVectorUniform", "acDeviceLoadIntUniform",
    "acDeviceLoadBoolUniform", "acDeviceLoadInt3Uniform", "acDeviceStoreScalarUniform",
    "acDeviceStoreVectorUniform", "acDeviceLoadMeshInfo", "acDeviceLoadStencilsFromConfig",
    "acDeviceLoadVertexBufferWithOffset",
    "acDeviceLoadMeshWithOffset", "acDeviceLoadVertexBuffer", "acDeviceLoadMesh",
    "acDeviceSetVertexBuffer", "acDeviceFlushOutputBuffers", "acDeviceStoreVertexBufferWithOffset",
    "acDeviceGetConfig", "acDeviceGetLocalConfig", "acDeviceStoreMeshWithOffset",
    "acDeviceStoreVertexBuffer", "acDeviceStoreMesh", "acDeviceIntegrateSubstep",
    "acDevicePeriodicBoundcondStep", "acDevicePeriodicBoundconds",
    "acDeviceReduceScalNoPostProcessing", "acDeviceReduceScal", "acDeviceReduceVecNoPostProcessing",
    "acDeviceReduceVec", "acDeviceReduceVecScalNoPostProcessing", "acDeviceReduceVecScal",
    "acDeviceUpdateTimestepCFL", "acDeviceGetTimestep", "acDeviceIntegrateSubstepDeviceDt",
    "acDeviceIntegrateStepsCFL", "acDeviceReduceDiagnostics", "acDeviceAutotune",
    "acDeviceGetSolveLaunchShape", "acDeviceGetSolveKernel", "acDeviceGetVertexBufferPtrs",
    "acDeviceSetSolveTiming",
    "acDeviceGetSolveTiming", "acDeviceGetStreamHandle",
    "acHostUpdateParams", "acSetGridMeshDims", "acSetLocalMeshDims", "acLoadConfig",
    "acSetDefaultConfig", "acHostMeshCreate", "acHostMeshDestroy", "acHostMeshCopyVertexBuffers",
    "acHostMeshCopy", "acHostMeshRandomize", "acHostMeshApplyPeriodicBounds",
    "acHostMeshInitFourier", "acGetRealSize", "acGetMeshInfoSize", "acGetMeshSize",
    "acGetBuildInfo",
    "acAdjointWorkspaceBytes", "acDeviceIntegrateSubstepTangent", "acDeviceIntegrateSubstepAdjoint",
    "acDeviceLoadMeshFromDevicePtr", "acDeviceStoreMeshToDevicePtr", "acDeviceInitFourier",
    "acHostAppendTimeseries", "acGetVertexBufferName", "acDeviceSetTimestepMode",
    "acDeviceWriteSnapshot", "acDeviceReadSnapshot", "acHostAppendSnapshotInfo",
    "acHostReadSnapshotInfo", "seed_rng", "random_uniform_real_01", "generateHelicalForcingParams",
    "loadForcingParamsToMeshInfo", "printForcingParams", "acDeviceLoadForcingParams",
    "acReduceNoAlloc",
]


class Int3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int), ("z", ctypes.c_int)]


class Volume(ctypes.Structure):
    _fields_ = [("x", ctypes.c_size_t), ("y", ctypes.c_size_t), ("z", ctypes.c_size_t)]


class AcDiagnostics(ctypes.Structure):
    _fields_ = [("min", ctypes.c_double * NUM_VTXBUF_HANDLES),
                ("max", ctypes.c_double * NUM_VTXBUF_HANDLES),
                ("rms", ctypes.c_double * NUM_VTXBUF_HANDLES),
                ("uu_rms", ctypes.c_double), ("uu_max", ctypes.c_double),
                ("bb_rms", ctypes.c_double), ("bb_max", ctypes.c_double),
                ("uu_min", ctypes.c_double), ("bb_min", ctypes.c_double),
                ("va_min", ctypes.c_double), ("va_rms", ctypes.c_double),
                ("va_max", ctypes.c_double)]


class AcReduction(ctypes.Structure):
    _fields_ = [("reduce_op", ctypes.c_int), ("post_processing_op", ctypes.c_int),
                ("map_vtxbuf_single", ctypes.c_int), ("map_vtxbuf_vec", ctypes.c_int),
                ("map_vtxbuf_vec_scal", ctypes.c_int), ("name", ctypes.c_char_p)]


def rtype(name):
    op, post, ms, mv, mvs = RTYPES[name]
    return AcReduction(op, post, ms, mv, mvs, name.encode())


def _make_types(c_real):
    class Real3(ctypes.Structure):
        _fields_ = [("x", c_real), ("y", c_real), ("z", c_real)]

    class AcMeshInfo(ctypes.Structure):
        _fields_ = [("int_params", ctypes.c_int * len(INT_PARAMS)),
                    ("int3_params", Int3 * len(INT3_PARAMS)),
                    ("real_params", c_real * len(REAL_PARAMS)),
                    ("real3_params", Real3 * len(REAL3_PARAMS)),
                    ("bool_params", ctypes.c_int * len(BOOL_PARAMS))]

        # dict-like access by parameter name, e.g. info["AC_nu_visc"] = 5e-3
        def __getitem__(self, name):
            if name in REAL_PARAMS:
                return self.real_params[REAL_PARAMS.index(name)]
            if name in INT_PARAMS:
                return self.int_params[INT_PARAMS.index(name)]
            if name in BOOL_PARAMS:
                return bool(self.bool_params[BOOL_PARAMS.index(name)])
            if name in INT3_PARAMS:
                v = self.int3_params[INT3_PARAMS.index(name)]
                return (v.x, v.y, v.z)
            if name in REAL3_PARAMS:
                v = self.real3_params[REAL3_PARAMS.index(name)]
                return (v.x, v.y, v.z)
            raise KeyError(name)

        def __setitem__(self, name, value):
            if name in REAL_PARAMS:
                self.real_params[REAL_PARAMS.index(name)] = value
            elif name in INT_PARAMS:
                self.int_params[INT_PARAMS.index(name)] = int(value)
            elif name in BOOL_PARAMS:
                self.bool_params[BOOL_PARAMS.index(name)] = 1 if value else 0
            elif name in INT3_PARAMS:
                self.int3_params[INT3_PARAMS.index(name)] = Int3(*[int(v) for v in value])
            elif name in REAL3_PARAMS:
                self.real3_params[REAL3_PARAMS.index(name)] = Real3(*value)
            else:
                raise KeyError(name)

    class AcMesh(ctypes.Structure):
        _fields_ = [("vertex_buffer", ctypes.POINTER(c_real) * NUM_VTXBUF_HANDLES),
                    ("info", AcMeshInfo)]

    return Real3, AcMeshInfo, AcMesh


def _bind_hip_runtime():
    """PyTorch-ROCm ships its own libamdhip64 (SONAME libamdhip64.so.7) and loads it by its
    unversioned file name. If this library were loaded first, the dynamic linker would bind
    /opt/rocm's runtime to it and torch would later load a second HIP runtime into the process,
    which finds no GPU. Importing torch first makes the library's libamdhip64.so.7 dependency
    resolve to the runtime torch already holds: one runtime per process, so device pointers
    and streams pass freely between torch tensors and this library. Without torch installed
    the system runtime is used."""
    if os.environ.get("ASTAROTH_SYSTEM_HIP_RUNTIME") == "1":
        return
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


class AstarothLib:
    """One precision build of the library (AcReal = float for fp32, double for fp64)."""

    def __init__(self, precision="fp32"):
        if precision not in ("fp32", "fp64"):
            raise ValueError(precision)
        path = os.path.join(LIBDIR, f"libastaroth_{precision}.so")
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing: build the HIP libraries first (python -m astaroth_amd.build)")
        self.path = path
        self.precision = precision
        _bind_hip_runtime()
        self.lib = ctypes.CDLL(path)
        self.c_real = ctypes.c_float if precision == "fp32" else ctypes.c_double
        self.np_real = np.float32 if precision == "fp32" else np.float64
        self.Real3, self.AcMeshInfo, self.AcMesh = _make_types(self.c_real)
        R3_ = self.Real3

        class ForcingParams(ctypes.Structure):
            """ref include/astaroth_forcing.h ForcingParams"""
            _fields_ = [("magnitude", self.c_real), ("k_force", R3_), ("ff_hel_re", R3_),
                        ("ff_hel_im", R3_), ("phase", self.c_real), ("kaver", self.c_real)]

        self.ForcingParams = ForcingParams
        if self.lib.acGetRealSize() != ctypes.sizeof(self.c_real):
            raise RuntimeError("AcReal size mismatch between library and binding")
        self._declare()

    def _declare(self):
        L, R = self.lib, self.c_real
        Info, Mesh, R3 = self.AcMeshInfo, self.AcMesh, self.Real3
        dev, i, vol = ctypes.c_void_p, ctypes.c_int, Volume
        P = ctypes.POINTER
        sig = {
            "acDeviceCreate": [i, Info, P(dev)],
            "acDeviceDestroy": [P(dev)],
            "acDeviceResetMesh": [dev, i],
            "acDevicePrintInfo": [dev],
            "acDeviceSynchronizeStream": [dev, i],
            "acDeviceSwapBuffer": [dev, i],
            "acDeviceSwapBuffers": [dev],
            "acDeviceLoadScalarUniform": [dev, i, i, R],
            "acDeviceLoadVectorUniform": [dev, i, i, R3],
            "acDeviceLoadIntUniform": [dev, i, i, i],
            "acDeviceLoadBoolUniform": [dev, i, i, ctypes.c_bool],
            "acDeviceLoadInt3Uniform": [dev, i, i, Int3],
            "acDeviceStoreScalarUniform": [dev, i, i, P(R)],
            "acDeviceStoreVectorUniform": [dev, i, i, P(R3)],
            "acDeviceLoadMeshInfo": [dev, Info],
            "acDeviceLoadStencilsFromConfig": [dev, i],
            "acDeviceLoadVertexBufferWithOffset": [dev, i, Mesh, i, Int3, Int3, i],
            "acDeviceLoadMeshWithOffset": [dev, i, Mesh, Int3, Int3, i],
            "acDeviceLoadVertexBuffer": [dev, i, Mesh, i],
            "acDeviceLoadMesh": [dev, i, Mesh],
            "acDeviceSetVertexBuffer": [dev, i, i, R],
            "acDeviceFlushOutputBuffers": [dev, i],
            "acDeviceStoreVertexBufferWithOffset": [dev, i, i, Int3, Int3, i, P(Mesh)],
            "acDeviceStoreMeshWithOffset": [dev, i, Int3, Int3, i, P(Mesh)],
            "acDeviceStoreVertexBuffer": [dev, i, i, P(Mesh)],
            "acDeviceStoreMesh": [dev, i, P(Mesh)],
            "acDeviceIntegrateSubstep": [dev, i, i, vol, vol, R],
            "acDevicePeriodicBoundcondStep": [dev, i, i, vol, vol],
            "acDevicePeriodicBoundconds": [dev, i, vol, vol],
            "acDeviceReduceScalNoPostProcessing": [dev, i, AcReduction, i, P(R)],
            "acDeviceReduceScal": [dev, i, AcReduction, i, P(R)],
            "acDeviceReduceVecNoPostProcessing": [dev, i, AcReduction, i, i, i, P(R)],
            "acDeviceReduceVec": [dev, i, AcReduction, i, i, i, P(R)],
            "acDeviceReduceVecScalNoPostProcessing": [dev, i, AcReduction, i, i, i, i, P(R)],
            "acDeviceReduceVecScal": [dev, i, AcReduction, i, i, i, i, P(R)],
            "acDeviceUpdateTimestepCFL": [dev, i, R],
            "acDeviceGetTimestep": [dev, i, P(ctypes.c_double)],
            "acDeviceIntegrateSubstepDeviceDt": [dev, i, i, R],
            "acDeviceIntegrateStepsCFL": [dev, i, i, R, P(ctypes.c_double)],
            "acDeviceReduceDiagnostics": [dev, i, P(AcDiagnostics)],
            "acDeviceAutotune": [dev, i],
            "acDeviceGetSolveLaunchShape": [dev, P(i), P(i), P(i)],
            "acDeviceGetSolveKernel": [dev, P(i)],
            "acDeviceGetVertexBufferPtrs": [dev, i, P(ctypes.c_void_p), P(ctypes.c_void_p),
                                            P(ctypes.c_void_p)],
            "acDeviceSetSolveTiming": [dev, i],
            "acDeviceGetSolveTiming": [dev, P(ctypes.c_double), P(i)],
            "acDeviceGetStreamHandle": [dev, i, P(ctypes.c_void_p)],
            "acHostUpdateParams": [P(Info)],
            "acSetGridMeshDims": [ctypes.c_size_t] * 3 + [P(Info)],
            "acSetLocalMeshDims": [ctypes.c_size_t] * 3 + [P(Info)],
            "acLoadConfig": [ctypes.c_char_p, P(Info)],
            "acSetDefaultConfig": [P(Info)],
            "acHostMeshCreate": [Info, P(Mesh)],
            "acHostMeshDestroy": [P(Mesh)],
            "acHostMeshCopyVertexBuffers": [Mesh, Mesh],
            "acHostMeshCopy": [Mesh, P(Mesh)],
            "acHostMeshRandomize": [P(Mesh)],
            "acHostMeshApplyPeriodicBounds": [P(Mesh)],
            "acHostMeshInitFourier": [P(Mesh), ctypes.c_ulonglong, i, i, ctypes.c_double,
                                      ctypes.c_double],
            "acGetRealSize": [],
            "acAdjointWorkspaceBytes": [dev],
            "acDeviceIntegrateSubstepTangent": [dev, i, i, R, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_void_p],
            "acDeviceIntegrateSubstepAdjoint": [dev, i, i, R, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_void_p],
            "acDeviceLoadMeshFromDevicePtr": [dev, i, ctypes.c_void_p],
            "acDeviceStoreMeshToDevicePtr": [dev, i, ctypes.c_void_p],
            "acDeviceInitFourier": [dev, i, ctypes.c_ulonglong, i, i, ctypes.c_double,
                                    ctypes.c_double],
            "acDeviceSetTimestepMode": [dev, i],
            "acDeviceWriteSnapshot": [dev, i, ctypes.c_char_p, ctypes.c_char_p],
            "acDeviceReadSnapshot": [dev, i, ctypes.c_char_p, ctypes.c_char_p],
            "acHostAppendSnapshotInfo": [ctypes.c_char_p, i, i, ctypes.c_double, P(Info)],
            "acHostReadSnapshotInfo": [ctypes.c_char_p, P(i), P(i), P(ctypes.c_double)],
            "loadForcingParamsToMeshInfo": [self.ForcingParams, P(Info)],
            "acDeviceLoadForcingParams": [dev, i, self.ForcingParams],
            "acReduceNoAlloc": [AcReduction, P(ctypes.c_void_p), i, i, i, i, i, ctypes.c_void_p,
                                ctypes.c_void_p],
        }
        sig["acHostAppendTimeseries"] = [ctypes.c_char_p, i, ctypes.c_double, ctypes.c_double,
                                         ctypes.POINTER(AcDiagnostics), ctypes.POINTER(ctypes.c_int)]
        for name, args in sig.items():
            if not hasattr(L, name):
                continue
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
            setattr(self, name, fn)
        L.acGetVertexBufferName.argtypes = [i]
        L.acGetVertexBufferName.restype = ctypes.c_char_p
        if hasattr(L, "generateHelicalForcingParams"):
            L.seed_rng.argtypes = [ctypes.c_uint32]
            L.seed_rng.restype = None
            L.random_uniform_real_01.argtypes = []
            L.random_uniform_real_01.restype = R
            L.generateHelicalForcingParams.argtypes = [R, R, R, R]
            L.generateHelicalForcingParams.restype = self.ForcingParams
            L.printForcingParams.argtypes = [self.ForcingParams]
            L.printForcingParams.restype = None
            for name in ("seed_rng", "random_uniform_real_01", "generateHelicalForcingParams",
                         "printForcingParams"):
                setattr(self, name, getattr(L, name))
        for name in ("acDeviceGetConfig", "acDeviceGetLocalConfig"):
            fn = getattr(L, name)
            fn.argtypes = [dev]
            fn.restype = Info
            setattr(self, name, fn)
        L.acGetBuildInfo.restype = ctypes.c_char_p
        L.acGetBuildInfo.argtypes = []
        self.acGetBuildInfo = L.acGetBuildInfo
        if hasattr(L, "acAdjointWorkspaceBytes"):
            L.acAdjointWorkspaceBytes.restype = ctypes.c_size_t
            self.acAdjointWorkspaceBytes = L.acAdjointWorkspaceBytes

    # ---- conveniences ----
    def config(self, n=(32, 32, 32), **params):
        """Default config (ref config/astaroth.conf) with interior n and overrides by name."""
        info = self.AcMeshInfo()
        self.acSetDefaultConfig(ctypes.byref(info))
        for k, v in params.items():
            info[k] = v
        check(self.acSetGridMeshDims(n[0], n[1], n[2], ctypes.byref(info)), "acSetGridMeshDims")
        check(self.acSetLocalMeshDims(n[0], n[1], n[2], ctypes.byref(info)), "acSetLocalMeshDims")
        return info

    def interior(self, info):
        n0 = info["AC_nmin"]
        n1 = info["AC_nlocal_max"]
        return Volume(*n0), Volume(*n1)

    def full(self, info):
        return Volume(0, 0, 0), Volume(*info["AC_mlocal"])


def check(ret, what="call"):
    if ret != AC_SUCCESS:
        raise RuntimeError(f"{what} returned {ret}")
    return ret


class HostMesh:
    """An AcMesh whose vertex buffers are views of one numpy array (NF, mz, my, mx)."""

    def __init__(self, lib, info, data=None):
        self.lib = lib
        self.info = info
        mx, my, mz = info["AC_mlocal"]
        if data is None:
            data = np.zeros((NUM_VTXBUF_HANDLES, mz, my, mx), dtype=lib.np_real)
        data = np.ascontiguousarray(data, dtype=lib.np_real)
        assert data.shape == (NUM_VTXBUF_HANDLES, mz, my, mx), data.shape
        self.data = data
        self.struct = lib.AcMesh()
        self.struct.info = info
        P = ctypes.POINTER(lib.c_real)
        for q in range(NUM_VTXBUF_HANDLES):
            self.struct.vertex_buffer[q] = self.data[q].ctypes.data_as(P)


_LIBS = {}


def load(precision="fp32"):
    if precision not in _LIBS:
        _LIBS[precision] = AstarothLib(precision)
    return _LIBS[precision]


class Device:
    """RAII wrapper over acDeviceCreate / acDeviceDestroy."""

    def __init__(self, lib, info, device_id=0):
        self.lib = lib
        self.handle = ctypes.c_void_p()
        check(lib.acDeviceCreate(device_id, info, ctypes.byref(self.handle)), "acDeviceCreate")
        self.info = lib.acDeviceGetLocalConfig(self.handle)

    def close(self):
        if self.handle:
            self.lib.acDeviceDestroy(ctypes.byref(self.handle))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # thin helpers, all in terms of the C calls
    def load_mesh(self, mesh, stream=STREAM_DEFAULT):
        check(self.lib.acDeviceLoadMesh(self.handle, stream, mesh.struct), "acDeviceLoadMesh")

    def store_mesh(self, mesh, stream=STREAM_DEFAULT):
        check(self.lib.acDeviceStoreMesh(self.handle, stream, ctypes.byref(mesh.struct)),
              "acDeviceStoreMesh")
        check(self.lib.acDeviceSynchronizeStream(self.handle, stream), "sync")

    def periodic(self, stream=STREAM_DEFAULT):
        a, b = self.lib.full(self.info)
        check(self.lib.acDevicePeriodicBoundconds(self.handle, stream, a, b),
              "acDevicePeriodicBoundconds")

    def substep(self, step, dt, stream=STREAM_DEFAULT):
        a, b = self.lib.interior(self.info)
        check(self.lib.acDeviceIntegrateSubstep(self.handle, stream, step, a, b, dt),
              "acDeviceIntegrateSubstep")

    def swap(self):
        check(self.lib.acDeviceSwapBuffers(self.handle), "acDeviceSwapBuffers")

    def step(self, dt, stream=STREAM_DEFAULT):
        """One full RK3 step in the reference's call pattern (samples/devicetest/main.cc:98-103)."""
        for s in range(3):
            self.periodic(stream)
            self.substep(s, dt, stream)
            self.swap()

    def sync(self, stream=STREAM_DEFAULT):
        check(self.lib.acDeviceSynchronizeStream(self.handle, stream), "sync")

    def init_fourier(self, seed, nmodes, kmax, ampl, lnrho_perturb=0.0, stream=STREAM_DEFAULT):
        """acDeviceInitFourier: the seeded Fourier initial condition generated on the device."""
        check(self.lib.acDeviceInitFourier(self.handle, stream, seed, nmodes, kmax, ampl,
                                           lnrho_perturb), "acDeviceInitFourier")

    def reduce_scal(self, name, field, post=True):
        out = self.lib.c_real()
        fn = self.lib.acDeviceReduceScal if post else self.lib.acDeviceReduceScalNoPostProcessing
        check(fn(self.handle, STREAM_DEFAULT, rtype(name), field, ctypes.byref(out)), name)
        return out.value

    def reduce_vec(self, name, a, b, c, post=True):
        out = self.lib.c_real()
        fn = self.lib.acDeviceReduceVec if post else self.lib.acDeviceReduceVecNoPostProcessing
        check(fn(self.handle, STREAM_DEFAULT, rtype(name), a, b, c, ctypes.byref(out)), name)
        return out.value

    def reduce_vec_scal(self, name, a, b, c, d, post=True):
        out = self.lib.c_real()
        fn = (self.lib.acDeviceReduceVecScal if post
              else self.lib.acDeviceReduceVecScalNoPostProcessing)
        check(fn(self.handle, STREAM_DEFAULT, rtype(name), a, b, c, d, ctypes.byref(out)), name)
        return out.value

    def diagnostics(self):
        d = AcDiagnostics()
        check(self.lib.acDeviceReduceDiagnostics(self.handle, STREAM_DEFAULT, ctypes.byref(d)),
              "acDeviceReduceDiagnostics")
        return d

    def stream_handle(self, stream=STREAM_DEFAULT):
        p = ctypes.c_void_p()
        check(self.lib.acDeviceGetStreamHandle(self.handle, stream, ctypes.byref(p)), "stream")
        return p.value

    def field_ptrs(self, field):
        a, b, c = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(self.lib.acDeviceGetVertexBufferPtrs(self.handle, field, ctypes.byref(a),
                                                   ctypes.byref(b), ctypes.byref(c)), "ptrs")
        return a.value, b.value, c.value
