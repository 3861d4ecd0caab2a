/*
 * astaroth.h -- C ABI of the MI355X-native single-device MHD hot path.
 *
 * Source-compatible subset of Astaroth's public device API for the
 * 6th-order MHD solve, the 2N-storage RK3 substep, the periodic halo fill and
 * the device reductions. Every function listed here is exported with C
 * linkage by libastaroth_fp32.so (AcReal = float) and libastaroth_fp64.so
 * (AcReal = double); the build picks the precision exactly as the reference's
 * DOUBLE_PRECISION option does (ref include/acreal.h, CMakeLists.txt:32-62).
 *
 * Reference interfaces replaced (file:line in the reference tree):
 *   types       include/astaroth_base.h:43-86,134-135, include/host_datatypes.h:10,
 *               acc-runtime/built-in/typedefs.h:218-222, acc-runtime/api/acc_runtime.h:92-100,157-175
 *   device API  include/astaroth_device.h:13-246 (implemented in src/core/device.cc)
 *   host API    include/astaroth_helpers.h:291-327, include/astaroth_utils.h:47-83
 *
 * Parameters keep the reference's names (AC_nu_visc, AC_ds, AC_nlocal, ...),
 * stored in fixed enum-indexed arrays. In the reference AcMeshInfo is generated
 * by the DSL compiler and has no stable binary layout, so compatibility is at
 * the source level: same function names, argument meaning and error codes.
 */
#ifndef ASTAROTH_AMD_H
#define ASTAROTH_AMD_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#define AC_BEGIN_C_DECLARATIONS extern "C" {
#define AC_END_C_DECLARATIONS }
#else
#define AC_BEGIN_C_DECLARATIONS
#define AC_END_C_DECLARATIONS
#endif

#ifndef AC_DOUBLE_PRECISION
#define AC_DOUBLE_PRECISION 0
#endif

#if AC_DOUBLE_PRECISION
typedef double AcReal;
#define AC_REAL_EPSILON (2.2204460492503131e-16)
#else
typedef float AcReal;
#define AC_REAL_EPSILON (1.19209290e-07f)
#endif

#define STENCIL_ORDER (6)
#define NGHOST (STENCIL_ORDER / 2)

/* ---- basic types (ref typedefs.h:7-12,218-222; host_datatypes.h:10) ---- */
#if defined(__HIP__)
#include <hip/hip_runtime.h> /* HIP's int3 (12 bytes, 4-byte aligned, same layout as below) */
#elif !defined(AC_HAVE_INT3)
#define AC_HAVE_INT3
typedef struct {
    int x, y, z;
} int3;
#endif

typedef struct {
    AcReal x, y, z;
} AcReal3;

typedef struct {
    size_t x, y, z;
} Volume;

typedef enum { AC_SUCCESS = 0, AC_FAILURE = 1, AC_NOT_ALLOCATED = 2 } AcResult;

/* ---- fields (ref mhdsolver.ac:125,136,140: declaration order) ---- */
typedef enum {
    VTXBUF_LNRHO = 0,
    VTXBUF_UUX,
    VTXBUF_UUY,
    VTXBUF_UUZ,
    VTXBUF_AX,
    VTXBUF_AY,
    VTXBUF_AZ,
    VTXBUF_ENTROPY,
    NUM_VTXBUF_HANDLES
} VertexBufferHandle;
typedef VertexBufferHandle Field;
#define NUM_FIELDS (NUM_VTXBUF_HANDLES)
#define NUM_ALL_FIELDS (NUM_VTXBUF_HANDLES)

/* ---- parameters (ref built-in/variables.h, mhdsolver.ac:196-222, config/astaroth.conf) ---- */
typedef enum {
    AC_max_steps = 0,
    AC_save_steps,
    AC_start_step,
    AC_kahan_bits,       /* FP32 build: bits of stored Kahan compensation per value (0 = off, 8, 16, 32) */
    AC_solve_tile_x,     /* launch shape of the solve kernel (0 = autotune) */
    AC_solve_tile_y,
    AC_solve_zchunk,
    AC_forcing_period_steps, /* steps between new forcing parameters (config/astaroth.conf:34) */
    AC_solve_kernel,     /* with a pinned launch shape: 0 = delay-line kernel, 1 = unrolled sweep (FP32) */
    NUM_INT_PARAMS
} AcIntParam;

typedef enum {
    AC_ngrid = 0,
    AC_nlocal,
    AC_mlocal,
    AC_mgrid,
    AC_nmin,
    AC_nlocal_max,
    AC_ngrid_max,
    AC_multigpu_offset,
    NUM_INT3_PARAMS
} AcInt3Param;

typedef enum {
    AC_cdt = 0,
    AC_cdtv,
    AC_cdts,
    AC_nu_visc,
    AC_cs_sound,
    AC_zeta,
    AC_eta,
    AC_mu0,
    AC_chi,
    AC_cp_sound,
    AC_gamma,
    AC_lnT0,
    AC_lnrho0,
    AC_K_heatcond,
    AC_cs2_sound,
    AC_dsmin,
    AC_current_time,
    AC_dt,
    /* helical forcing (ref mhdsolver.ac:239-257, config/astaroth.conf:34-35,58-61) */
    AC_forcing_magnitude,
    AC_relhel,
    AC_kmin,
    AC_kmax,
    AC_forcing_phase,
    AC_k_forcex,
    AC_k_forcey,
    AC_k_forcez,
    AC_ff_hel_rex,
    AC_ff_hel_rey,
    AC_ff_hel_rez,
    AC_ff_hel_imx,
    AC_ff_hel_imy,
    AC_ff_hel_imz,
    AC_kaver,
    AC_switch_forcing,
    AC_forcing_period_t,
    NUM_REAL_PARAMS
} AcRealParam;

typedef enum {
    AC_ds = 0,
    AC_inv_ds,
    AC_inv_ds_2,
    AC_len,
    NUM_REAL3_PARAMS
} AcReal3Param;

typedef enum {
    AC_lupwd = 0,        /* ref hostdefine LUPWD (mhdsolver.ac:101): 6th-order upwind hyperdiffusion */
    AC_lentropy,         /* ref hostdefine LENTROPY (mhdsolver.ac:99): evolve VTXBUF_ENTROPY */
    AC_additive_timestep,
    AC_periodic_grid_all,
    AC_lforcing,         /* ref mhdsolver.ac:533,622: add the helical forcing to u at substep 2 */
    NUM_BOOL_PARAMS
} AcBoolParam;

typedef struct AcMeshInfo {
    int int_params[NUM_INT_PARAMS];
    int3 int3_params[NUM_INT3_PARAMS];
    AcReal real_params[NUM_REAL_PARAMS];
    AcReal3 real3_params[NUM_REAL3_PARAMS];
    int bool_params[NUM_BOOL_PARAMS];
#ifdef __cplusplus
    int& operator[](const AcIntParam p) { return int_params[p]; }
    const int& operator[](const AcIntParam p) const { return int_params[p]; }
    int3& operator[](const AcInt3Param p) { return int3_params[p]; }
    const int3& operator[](const AcInt3Param p) const { return int3_params[p]; }
    AcReal& operator[](const AcRealParam p) { return real_params[p]; }
    const AcReal& operator[](const AcRealParam p) const { return real_params[p]; }
    AcReal3& operator[](const AcReal3Param p) { return real3_params[p]; }
    const AcReal3& operator[](const AcReal3Param p) const { return real3_params[p]; }
    int& operator[](const AcBoolParam p) { return bool_params[p]; }
    const int& operator[](const AcBoolParam p) const { return bool_params[p]; }
#endif
} AcMeshInfo;

/* ---- host mesh (ref astaroth_base.h:43-47) ---- */
typedef struct {
    AcReal* vertex_buffer[NUM_ALL_FIELDS];
    AcMeshInfo info;
} AcMesh;

/* ---- streams (ref astaroth_base.h:49-86). All are created at default priority. ---- */
typedef enum {
    STREAM_0, STREAM_1, STREAM_2, STREAM_3, STREAM_4, STREAM_5, STREAM_6, STREAM_7,
    STREAM_8, STREAM_9, STREAM_10, STREAM_11, STREAM_12, STREAM_13, STREAM_14, STREAM_15,
    STREAM_16, STREAM_17, STREAM_18, STREAM_19, STREAM_20, STREAM_21, STREAM_22, STREAM_23,
    STREAM_24, STREAM_25, STREAM_26, STREAM_27, STREAM_28, STREAM_29, STREAM_30, STREAM_31,
    STREAM_ALL
} Stream;
#define STREAM_DEFAULT (STREAM_0)
#define NUM_STREAMS (32)

typedef struct device_s* Device;

/* ---- reductions (ref acc_runtime.h:92-100, stdlib/reduction.h) ---- */
typedef enum {
    AC_REDUCE_OP_NO_REDUCE = 0,
    AC_REDUCE_OP_MIN,
    AC_REDUCE_OP_MAX,
    AC_REDUCE_OP_SUM
} AcReduceOp;

typedef enum { AC_NO_REDUCE_POST_PROCESSING = 0, AC_RMS = 1 } AcReducePostProcessingOp;

typedef enum {
    AC_NULL_KERNEL = 0,
    AC_MAP_VTXBUF,
    AC_MAP_VTXBUF_SQUARE,
    AC_MAP_VTXBUF_EXP_SQUARE,
    AC_MAP_VTXBUF3_NORM,
    AC_MAP_VTXBUF3_SQUARE,
    AC_MAP_VTXBUF3_EXP_SQUARE,
    AC_MAP_VTXBUF4_ALFVEN_NORM,
    AC_MAP_VTXBUF4_ALFVEN_SQUARE
} AcMapKernel;

typedef struct {
    AcReduceOp reduce_op;
    AcReducePostProcessingOp post_processing_op;
    AcMapKernel map_vtxbuf_single;
    AcMapKernel map_vtxbuf_vec;
    AcMapKernel map_vtxbuf_vec_scal;
    const char* name;
} AcReduction;

static const AcReduction RTYPE_MAX = {AC_REDUCE_OP_MAX, AC_NO_REDUCE_POST_PROCESSING, AC_MAP_VTXBUF,
                                      AC_MAP_VTXBUF3_NORM, AC_NULL_KERNEL, "RTYPE_MAX"};
static const AcReduction RTYPE_MIN = {AC_REDUCE_OP_MIN, AC_NO_REDUCE_POST_PROCESSING, AC_MAP_VTXBUF,
                                      AC_MAP_VTXBUF3_NORM, AC_NULL_KERNEL, "RTYPE_MIN"};
static const AcReduction RTYPE_SUM = {AC_REDUCE_OP_SUM, AC_NO_REDUCE_POST_PROCESSING, AC_MAP_VTXBUF,
                                      AC_MAP_VTXBUF3_NORM, AC_NULL_KERNEL, "RTYPE_SUM"};
static const AcReduction RTYPE_RMS = {AC_REDUCE_OP_SUM, AC_RMS, AC_MAP_VTXBUF_SQUARE,
                                      AC_MAP_VTXBUF3_SQUARE, AC_NULL_KERNEL, "RTYPE_RMS"};
static const AcReduction RTYPE_RMS_EXP = {AC_REDUCE_OP_SUM, AC_RMS, AC_MAP_VTXBUF_EXP_SQUARE,
                                          AC_MAP_VTXBUF3_EXP_SQUARE, AC_NULL_KERNEL,
                                          "RTYPE_RMS_EXP"};
static const AcReduction RTYPE_ALFVEN_MAX = {AC_REDUCE_OP_MAX, AC_NO_REDUCE_POST_PROCESSING,
                                             AC_NULL_KERNEL, AC_NULL_KERNEL,
                                             AC_MAP_VTXBUF4_ALFVEN_NORM, "RTYPE_ALFVEN_MAX"};
static const AcReduction RTYPE_ALFVEN_MIN = {AC_REDUCE_OP_MIN, AC_NO_REDUCE_POST_PROCESSING,
                                             AC_NULL_KERNEL, AC_NULL_KERNEL,
                                             AC_MAP_VTXBUF4_ALFVEN_NORM, "RTYPE_ALFVEN_MIN"};
static const AcReduction RTYPE_ALFVEN_RMS = {AC_REDUCE_OP_SUM, AC_RMS, AC_NULL_KERNEL,
                                             AC_NULL_KERNEL, AC_MAP_VTXBUF4_ALFVEN_SQUARE,
                                             "RTYPE_ALFVEN_RMS"};

/* One-pass diagnostics of every field (no reference counterpart as one call; the same
 * quantities are what ref samples/standalone_mpi/main.cc:256-391 prints per step). */
typedef struct {
    double min[NUM_VTXBUF_HANDLES];
    double max[NUM_VTXBUF_HANDLES];
    double rms[NUM_VTXBUF_HANDLES];
    double uu_rms;   /* sqrt(<|u|^2>) */
    double uu_max;   /* max |u| */
    double bb_rms;   /* sqrt(<|curl A|^2>) */
    double bb_max;   /* max |curl A| */
    double uu_min;   /* min |u| */
    double bb_min;   /* min |curl A| */
    double va_min;   /* Alfven speed |curl A| / sqrt(exp(lnrho)): min, rms, max */
    double va_rms;
    double va_max;
} AcDiagnostics;

AC_BEGIN_C_DECLARATIONS

/* ===================== Device interface (ref include/astaroth_device.h) ===================== */

/* ref astaroth_device.h:13, device.cc:508-593 */
AcResult acDeviceCreate(const int id, const AcMeshInfo device_config, Device* device);
/* ref astaroth_device.h:15, device.cc:601-632 */
AcResult acDeviceDestroy(Device* device);
/* ref astaroth_device.h:18 */
AcResult acDeviceResetMesh(const Device device, const Stream stream);
/* ref astaroth_device.h:20 */
AcResult acDevicePrintInfo(const Device device);
/* ref astaroth_device.h:24 */
AcResult acDeviceSynchronizeStream(const Device device, const Stream stream);
/* ref astaroth_device.h:26, device.cc:634-651 */
AcResult acDeviceSwapBuffer(const Device device, const VertexBufferHandle handle);
/* ref astaroth_device.h:28, device.cc:653-663 */
AcResult acDeviceSwapBuffers(const Device device);
/* ref astaroth_device.h:32 */
AcResult acDeviceLoadScalarUniform(const Device device, const Stream stream,
                                   const AcRealParam param, const AcReal value);
/* ref astaroth_device.h:47 */
AcResult acDeviceLoadVectorUniform(const Device device, const Stream stream,
                                   const AcReal3Param param, const AcReal3 value);
/* ref astaroth_device.h:50 */
AcResult acDeviceLoadIntUniform(const Device device, const Stream stream, const AcIntParam param,
                                const int value);
/* ref astaroth_device.h:52 */
AcResult acDeviceLoadBoolUniform(const Device device, const Stream stream,
                                 const AcBoolParam param, const bool value);
/* ref astaroth_device.h:55 */
AcResult acDeviceLoadInt3Uniform(const Device device, const Stream stream,
                                 const AcInt3Param param, const int3 value);
/* ref astaroth_device.h:58 */
AcResult acDeviceStoreScalarUniform(const Device device, const Stream stream,
                                    const AcRealParam param, AcReal* value);
/* ref astaroth_device.h:61 */
AcResult acDeviceStoreVectorUniform(const Device device, const Stream stream,
                                    const AcReal3Param param, AcReal3* value);
/* ref astaroth_device.h:64, device.cc:394-413 */
AcResult acDeviceLoadMeshInfo(const Device device, const AcMeshInfo device_config);
/* ref astaroth_device.h:67, device.cc:665-686 */
AcResult acDeviceLoadVertexBufferWithOffset(const Device device, const Stream stream,
                                            const AcMesh host_mesh,
                                            const VertexBufferHandle vtxbuf_handle, const int3 src,
                                            const int3 dst, const int num_vertices);
/* ref astaroth_device.h:73 */
AcResult acDeviceLoadMeshWithOffset(const Device device, const Stream stream,
                                    const AcMesh host_mesh, const int3 src, const int3 dst,
                                    const int num_vertices);
/* ref astaroth_device.h:77, device.cc:702-720 */
AcResult acDeviceLoadVertexBuffer(const Device device, const Stream stream, const AcMesh host_mesh,
                                  const VertexBufferHandle vtxbuf_handle);
/* ref astaroth_device.h:80, device.cc:736-745 */
AcResult acDeviceLoadMesh(const Device device, const Stream stream, const AcMesh host_mesh);
/* ref astaroth_device.h:93 */
AcResult acDeviceSetVertexBuffer(const Device device, const Stream stream,
                                 const VertexBufferHandle handle, const AcReal value);
/* ref astaroth_device.h:96 */
AcResult acDeviceFlushOutputBuffers(const Device device, const Stream stream);
/* ref astaroth_device.h:98, device.cc:788-810 */
AcResult acDeviceStoreVertexBufferWithOffset(const Device device, const Stream stream,
                                             const VertexBufferHandle vtxbuf_handle,
                                             const int3 src, const int3 dst,
                                             const int num_vertices, AcMesh* host_mesh);
/* ref astaroth_device.h:102 */
AcMeshInfo acDeviceGetConfig(const Device device);
/* ref astaroth_device.h:246 */
AcMeshInfo acDeviceGetLocalConfig(const Device device);
/* ref astaroth_device.h:108 */
AcResult acDeviceStoreMeshWithOffset(const Device device, const Stream stream, const int3 src,
                                     const int3 dst, const int num_vertices, AcMesh* host_mesh);
/* ref astaroth_device.h:112, device.cc:826-843 */
AcResult acDeviceStoreVertexBuffer(const Device device, const Stream stream,
                                   const VertexBufferHandle vtxbuf_handle, AcMesh* host_mesh);
/* ref astaroth_device.h:115, device.cc:845-854 */
AcResult acDeviceStoreMesh(const Device device, const Stream stream, AcMesh* host_mesh);
/* ref astaroth_device.h:133, device.cc:940-996.
 * One fused launch: w <- alpha_s*w + dt*RHS(in), out <- in + beta_s*w (interior only).
 * After the call the new state is in `out`; the caller swaps, as with the reference. */
AcResult acDeviceIntegrateSubstep(const Device device, const Stream stream, const int step_number,
                                  const Volume start, const Volume end, const AcReal dt);
/* ref astaroth_device.h:135, device.cc:998-1007: the halo cells of one field inside the box
 * [start, end) (0 <= start <= end <= mlocal) take their periodic images; interior cells are
 * left alone. */
AcResult acDevicePeriodicBoundcondStep(const Device device, const Stream stream,
                                       const VertexBufferHandle vtxbuf_handle, const Volume start,
                                       const Volume end);
/* ref astaroth_device.h:229, device.cc:447-473: evaluates the stencil coefficients from the
 * device's AC_ds; they stay in use until the next call (acDeviceCreate makes the first). The
 * reference's acDeviceLoadStencil(s) / acDeviceStoreStencil take DSL-generated 7x7x7 stencil
 * tables whose enumeration is generated per DSL program; this build's stencils are the fixed
 * 6th-order set, so only this form is provided. */
AcResult acDeviceLoadStencilsFromConfig(const Device device, const Stream stream);
/* ref astaroth_device.h:139, device.cc:1009-1017: the same for every field, in one launch (the
 * full volume [0, mlocal) enumerates only its halo cells) */
AcResult acDevicePeriodicBoundconds(const Device device, const Stream stream, const Volume start,
                                    const Volume end);
/* ref astaroth_device.h:149 */
AcResult acDeviceReduceScalNoPostProcessing(const Device device, const Stream stream,
                                            const AcReduction reduction,
                                            const VertexBufferHandle vtxbuf_handle,
                                            AcReal* result);
/* ref astaroth_device.h:153, device.cc:1403-1413 */
AcResult acDeviceReduceScal(const Device device, const Stream stream, const AcReduction reduction,
                            const VertexBufferHandle vtxbuf_handle, AcReal* result);
/* ref astaroth_device.h:157 */
AcResult acDeviceReduceVecNoPostProcessing(const Device device, const Stream stream,
                                           const AcReduction reduction,
                                           const VertexBufferHandle vtxbuf0,
                                           const VertexBufferHandle vtxbuf1,
                                           const VertexBufferHandle vtxbuf2, AcReal* result);
/* ref astaroth_device.h:161, device.cc:1415-1428 */
AcResult acDeviceReduceVec(const Device device, const Stream stream, const AcReduction reduction,
                           const VertexBufferHandle vtxbuf0, const VertexBufferHandle vtxbuf1,
                           const VertexBufferHandle vtxbuf2, AcReal* result);
/* ref astaroth_device.h:166 */
AcResult acDeviceReduceVecScalNoPostProcessing(const Device device, const Stream stream,
                                               const AcReduction reduction,
                                               const VertexBufferHandle vtxbuf0,
                                               const VertexBufferHandle vtxbuf1,
                                               const VertexBufferHandle vtxbuf2,
                                               const VertexBufferHandle vtxbuf3, AcReal* result);
/* ref astaroth_device.h:172, device.cc:1430-1443 */
AcResult acDeviceReduceVecScal(const Device device, const Stream stream,
                               const AcReduction reduction, const VertexBufferHandle vtxbuf0,
                               const VertexBufferHandle vtxbuf1, const VertexBufferHandle vtxbuf2,
                               const VertexBufferHandle vtxbuf3, AcReal* result);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the reduction of
 * acDeviceReduceScal (num_vtxbufs = 1), acDeviceReduceVec (3) or acDeviceReduceVecScal (4), with
 * (post_process != 0) or without the AC_RMS post-processing, over caller-owned device buffers of
 * (nx + 6) (ny + 6) (nz + 6) values each (the halo'ed layout of the device's vertex buffers,
 * x fastest; only the interior is reduced), enqueued on the HIP stream `stream` (0: the null
 * stream). The AcReal result is written to the device address `result`; nothing synchronises
 * the host. Its partial results live in a fixed device array of 1025 doubles (8 KiB per GPU,
 * held until the process ends), which bounds the first stage at 1024 workgroups for every grid
 * size; no allocation is made. The same kernels as the allocating entry points: the same value
 * bit for bit. Calls that overlap on one GPU share that array and race. Returns AC_FAILURE
 * (nothing launched) for null pointers, non-positive extents or a reduction type undefined for
 * the number of buffers. */
AcResult acReduceNoAlloc(const AcReduction reduction, const AcReal* const* vtxbufs,
                         const int num_vtxbufs, const int nx, const int ny, const int nz,
                         const int post_process, void* stream, AcReal* result);

/* ---- Extensions with no one-call counterpart in the reference ---- */

/* Timestep formula of the device-resident dt below.
 * AC_TIMESTEP_COURANT (default): dt = courant * dsmin / (cs_sound + max|u|), max|u| of the state
 *   produced by the last step.
 * AC_TIMESTEP_REFERENCE: the reference's CFL (ref samples/standalone_mpi/main.cc:395-440 and
 *   acc-runtime/samples/mhd_modular/mhdsolver.ac:551-583): with AC_additive_timestep false,
 *   dt = min(AC_cdt dsmin / (max|u| + sqrt(cs^2 + max vA^2)), AC_cdtv dsmin^2 / max(nu, eta));
 *   with it true, dt = min over cells of AC_cdt dsmin / (|u| + sqrt(cs^2 + vA^2)) and
 *   AC_cdtv dsmin^2 / max(nu, eta, gamma chi); vA = |curl A| / sqrt(exp(lnrho)). The maxima are
 *   reduced over the state entering the third substep, inside that substep, as the reference
 *   does; the courant arguments below are then ignored. */
enum { AC_TIMESTEP_COURANT = 0, AC_TIMESTEP_REFERENCE = 1 };
AcResult acDeviceSetTimestepMode(const Device device, const int mode);
/* Device-resident CFL timestep, reduced on the device in FP64 from the current state. Replaces
 * the host round trip of ref samples/standalone_mpi/main.cc:395-440 (which reads UU_MAX_ADVEC
 * back each step). */
AcResult acDeviceUpdateTimestepCFL(const Device device, const Stream stream, const AcReal courant);
/* Returns the current device-resident dt (synchronises the stream). */
AcResult acDeviceGetTimestep(const Device device, const Stream stream, double* dt);
/* As acDeviceIntegrateSubstep, with dt read on the device from the value the last
 * acDeviceUpdateTimestepCFL (or step_number == 2 of this call) produced. At step_number == 2
 * the kernel also reduces max|u| of the new state and sets the next dt (courant as given). */
AcResult acDeviceIntegrateSubstepDeviceDt(const Device device, const Stream stream,
                                          const int step_number, const AcReal courant);
/* nsteps full RK3 steps (periodic boundaries + fused substep + swap, three times per step) with
 * the device-resident CFL dt; no host synchronisation inside. The tiled solve kernels read the
 * periodic images of the halo cells themselves (no separate fill before their substeps; other
 * kernels get one), so the result equals fill + substep + swap bit for bit, and the halos of the
 * final state are not refreshed (apply acDevicePeriodicBoundconds before reading them).
 * Accumulated simulated time is added to *time_out when it is not NULL (synchronises at the end). */
AcResult acDeviceIntegrateStepsCFL(const Device device, const Stream stream, const int nsteps,
                                   const AcReal courant, double* time_out);
/* One pass over the interior: per-field min/max/rms, |u| and |curl A| rms/max (FP64). */
AcResult acDeviceReduceDiagnostics(const Device device, const Stream stream, AcDiagnostics* out);
/* Runs the launch autotuner of the solve kernel for the current grid now (it also runs on the
 * first substep). Writes the chosen shape into the cache file named by
 * ASTAROTH_AUTOTUNE_CACHE (default: no file). */
AcResult acDeviceAutotune(const Device device, const Stream stream);
/* Reports the solve-kernel launch shape in use (tile_x, tile_y, zchunk). */
AcResult acDeviceGetSolveLaunchShape(const Device device, int* tile_x, int* tile_y, int* zchunk);
/* Reports which solve kernel that shape runs: -1 direct, 0 delay-line, 1 unrolled sweep. */
AcResult acDeviceGetSolveKernel(const Device device, int* kind);
/* Raw device pointers of one field: in (current state), out (next state), w (RK register). */
AcResult acDeviceGetVertexBufferPtrs(const Device device, const VertexBufferHandle vtxbuf,
                                     AcReal** in, AcReal** out, AcReal** w);
/* Solve-kernel timing with HIP events recorded around every solve launch on the stream it is
 * launched on. enable != 0 starts a new measurement window (up to 4096 launches). */
AcResult acDeviceSetSolveTiming(const Device device, const int enable);
/* Synchronises the device and reports the summed device time (ms) and number of solve
 * launches recorded since the window started. */
AcResult acDeviceGetSolveTiming(const Device device, double* total_ms, int* launches);
/* The hipStream_t behind a Stream (for interop with other HIP code; default priority). */
AcResult acDeviceGetStreamHandle(const Device device, const Stream stream, void** hip_stream);

/* Copies the whole state (all fields, m^3 each, field q at src + q m^3) from / to device memory
 * the caller owns (interop with other device code; no reference counterpart). */
AcResult acDeviceLoadMeshFromDevicePtr(const Device device, const Stream stream, const AcReal* src);
AcResult acDeviceStoreMeshToDevicePtr(const Device device, const Stream stream, AcReal* dst);

/* Snapshots in the reference's on-disk format (ref src/core/grid.cc:4405-4530, collective IO on
 * one device): per evolved field one raw file dir/<field name>-<label>.mesh of the interior
 * cells (nx*ny*nz AcReal, x fastest). Write stores the current state (in); read loads it into
 * the current state, fills the halos periodically and clears the compensation; a file whose
 * size does not match the grid is an error. Both synchronise the stream. */
AcResult acDeviceWriteSnapshot(const Device device, const Stream stream, const char* dir,
                               const char* label);
AcResult acDeviceReadSnapshot(const Device device, const Stream stream, const char* dir,
                              const char* label);

/* The seeded Fourier-mode initial condition of acHostMeshInitFourier, generated directly in the
 * device's input buffers (bit-identical to the host version followed by acDeviceLoadMesh), with
 * halos filled and the output and compensation buffers cleared. For grids whose host mesh is
 * impractical to build (1024^3 and up). No reference counterpart. Synchronises the stream. */
AcResult acDeviceInitFourier(const Device device, const Stream stream, const unsigned long long seed,
                             const int nmodes, const int kmax, const double ampl,
                             const double lnrho_perturb);

/* ---- Differentiation of one RK3 substep (no reference counterpart) ----
 * Linearisation point: the device's current state `in` (halos filled by the caller, as before
 * acDeviceIntegrateSubstep). Slabs hold NUM_VTXBUF_HANDLES fields of m^3 values each.
 * Tangent:  dw' = alpha_s dw + dt (J df + dR/dnu dnu + dR/deta deta),  df' = df + beta_s dw'.
 *   df_in: tangent of f (its halos are filled by the call), df_out: df', dw: updated in place,
 *   dparams: host pointer to {dnu, deta} (NULL = zero).
 * Adjoint:  mu = lw + beta_s lf,  lf <- lf + dt J^T mu,  lw <- alpha_s mu (both in place);
 *   param_grad: device pointer to two doubles, += dt <mu, dR/dnu>, dt <mu, dR/deta>.
 *   workspace: device memory of acAdjointWorkspaceBytes(device) bytes. */
size_t acAdjointWorkspaceBytes(const Device device);
AcResult acDeviceIntegrateSubstepTangent(const Device device, const Stream stream,
                                         const int step_number, const AcReal dt,
                                         const AcReal* df_in, AcReal* df_out, AcReal* dw,
                                         const double* dparams);
AcResult acDeviceIntegrateSubstepAdjoint(const Device device, const Stream stream,
                                         const int step_number, const AcReal dt, AcReal* lam_f,
                                         AcReal* lam_w, void* workspace, double* param_grad);

/* ===================== Host interface (ref include/astaroth_helpers.h, astaroth_utils.h) ===== */

/* ref astaroth_helpers.h:291, src/core/update_params/update_params.cc:19-142 */
AcResult acHostUpdateParams(AcMeshInfo* config);
/* ref astaroth_helpers.h:322 */
AcResult acSetGridMeshDims(const size_t nx, const size_t ny, const size_t nz, AcMeshInfo* info);
/* ref astaroth_helpers.h:327 */
AcResult acSetLocalMeshDims(const size_t nx, const size_t ny, const size_t nz, AcMeshInfo* info);
/* ref astaroth_utils.h:47-83, src/utils/config_loader.cc:198-315: "AC_name = value" lines */
AcResult acLoadConfig(const char* config_path, AcMeshInfo* config);
/* Fills a config with the defaults of ref config/astaroth.conf (no file needed). */
AcResult acSetDefaultConfig(AcMeshInfo* config);
/* ref astaroth_helpers.h:305, src/core/astaroth.cc:295-303 */
AcResult acHostMeshCreate(const AcMeshInfo mesh_info, AcMesh* mesh);
/* ref astaroth_helpers.h:319 */
AcResult acHostMeshDestroy(AcMesh* mesh);
/* ref astaroth_helpers.h:307 */
AcResult acHostMeshCopyVertexBuffers(const AcMesh src, AcMesh dst);
/* ref astaroth_helpers.h:309 */
AcResult acHostMeshCopy(const AcMesh src, AcMesh* dst);
/* ref astaroth_helpers.h:315, src/core/astaroth.cc:382-395: rand()/RAND_MAX per value */
AcResult acHostMeshRandomize(AcMesh* mesh);
/* ref src/utils/memory.cc:75-149 */
AcResult acHostMeshApplyPeriodicBounds(AcMesh* mesh);
/* Seeded solenoidal Fourier-mode initial condition (no reference counterpart; the reference's
 * init types are ref astaroth_base.h:88-99). lnrho = ss = 0 (+ uniform perturbation of
 * amplitude lnrho_perturb in [-a, a) when nonzero); uu and aa each a sum of nmodes modes
 * ampl * e * cos(k.x + phi), integer k with 1 <= |k| <= kmax, e a unit vector normal to k.
 * Uses SplitMix64 (seed, seed+1, seed+2 for uu, aa, lnrho). FP64 arithmetic, cast to AcReal. */
AcResult acHostMeshInitFourier(AcMesh* mesh, const unsigned long long seed, const int nmodes,
                               const int kmax, const double ampl, const double lnrho_perturb);

/* Name of a field as the reference prints it ("VTXBUF_LNRHO", ...; ref vtxbuf_names). */
const char* acGetVertexBufferName(const VertexBufferHandle handle);

/* Appends one row of diagnostics to a timeseries file in the format of the reference's
 * standalone driver (ref samples/standalone_mpi/main.cc:218-391, timeseries.ts), writing the
 * header first when the file is new or empty. *found_nan (if given) is set when a per-field
 * reduction is NaN, the condition on which that driver stops. */
AcResult acHostAppendTimeseries(const char* path, const int step, const double t_step,
                                const double dt, const AcDiagnostics* diag, int* found_nan);

/* Snapshot restart table in the reference's format (ref samples/standalone_mpi/main.cc:186-203):
 * dir/snapshots_info.csv, truncated with a header at step 0, then one row per snapshot
 * "use_double,mx,my,mz,step_number,modstep,t_step,t_step(exact)". */
AcResult acHostAppendSnapshotInfo(const char* dir, const int step, const int modstep,
                                  const double t_step, const AcMeshInfo* info);
/* The last row of dir/snapshots_info.csv (the snapshot a restart loads, ref main.cc:579-605). */
AcResult acHostReadSnapshotInfo(const char* dir, int* step, int* modstep, double* t_step);

/* ---- Helical forcing (ref include/astaroth_forcing.h, include/astaroth_random.h) ----
 * With AC_lforcing set, the last substep of every RK3 step adds to u the forcing of
 * ref mhdsolver.ac:286-331,614-631 (while AC_current_time > AC_switch_forcing):
 *   f(x) = sqrt(dt) cs AC_forcing_magnitude sqrt(AC_kaver cs)
 *          (ff_re cos(k.y + phase) - ff_im sin(k.y + phase)),  y = 2 pi (x - x0) / (ds n),
 * with k = (AC_k_forcex, ...), ff_re = (AC_ff_hel_rex, ...), ff_im = (AC_ff_hel_imx, ...). The
 * parameters come from generateHelicalForcingParams and reach a device through
 * loadForcingParamsToMeshInfo + acDeviceLoadMeshInfo or acDeviceLoadForcingParams. */
typedef struct {
    AcReal magnitude;
    AcReal3 k_force;
    AcReal3 ff_hel_re;
    AcReal3 ff_hel_im;
    AcReal phase;
    AcReal kaver;
} ForcingParams;

/* Seeds the process-wide std::mt19937 the forcing generator draws from (ref
 * src/utils/astaroth_random.cc:seed_rng; initial seed 19937123 as there). */
void seed_rng(uint32_t seed);
/* One uniform [0, 1) draw from that generator (ref astaroth_random.cc:random_uniform_real_01). */
AcReal random_uniform_real_01(void);
/* A random helical forcing (ref src/utils/helical_forcing.cc:generateHelicalForcingParams):
 * integer k with kmin <= |k| <= kmax, drawn uniformly from that shell; a random unit e not
 * parallel to k; f_re = |k| (k x e) / D and f_im = relhel k x (k x e) / D with
 * D = sqrt(1 + relhel^2) |k|^2 sqrt(1 - (k.e)^2/|k|^2), so |f_re|^2 + |f_im|^2 = 1; phase uniform
 * in [-pi, pi); kaver = (kmax - kmin) / 2 + kmin. Same draws, in the same order, from the same
 * generator as the reference, so a seeded sequence of parameters matches it. */
ForcingParams generateHelicalForcingParams(const AcReal relhel, const AcReal magnitude,
                                           const AcReal kmin, const AcReal kmax);
/* Copies the parameters into the forcing fields of *info (ref
 * src/utils/astaroth_forcing.cc:loadForcingParamsToMeshInfo). Returns 0. */
int loadForcingParamsToMeshInfo(const ForcingParams forcing_params, AcMeshInfo* info);
/* Prints the parameters to stdout (ref astaroth_forcing.cc:printForcingParams). */
void printForcingParams(const ForcingParams forcing_params);
/* The forcing fields of the device's configuration, as the reference's driver loads them with
 * acGridLoadScalarUniform (ref samples/standalone_mpi/host_forcing.cc:loadForcingParamsToGrid). */
AcResult acDeviceLoadForcingParams(const Device device, const Stream stream,
                                   const ForcingParams forcing_params);

/* Library identification: sizeof(AcReal), sizeof(AcMeshInfo), sizeof(AcMesh) of this build and
 * the offload target it was built for (a caller checks its own struct sizes against these, as
 * ref src/core/astaroth.cc acVerifyCompatibility does). */
int acGetRealSize(void);
size_t acGetMeshInfoSize(void);
size_t acGetMeshSize(void);
const char* acGetBuildInfo(void);

AC_END_C_DECLARATIONS

#endif /* ASTAROTH_AMD_H */
