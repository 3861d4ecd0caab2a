// device.hip -- the device layer of the C API (ref src/core/device.cc, include/astaroth_device.h)
// on HIP for gfx950.
//
// Memory: one slab per role (in, out, w, and the optional Kahan compensation), each holding
// NUM_VTXBUF_HANDLES contiguous m^3 arrays in the reference's x-fastest layout
// (ref acc-runtime/api/vba.cc:313-470 allocates the in/out slabs the same way).
// acDeviceSwapBuffer(s) swaps the in/out pointers of fields; w and the compensation stay put
// because they are updated in place, point by point.
//
// Error behaviour: the reference aborts (ERRCHK_ALWAYS) on invariant violations such as a
// device id out of range, a host/device size mismatch or a partial integration volume. Here the
// same conditions print the reason to stderr and return AC_FAILURE, so a caller (or a test) can
// recover; AC_NOT_ALLOCATED keeps its meaning for absent host buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/astaroth.h"
#include "kernels.h"

using namespace acamd;

#define AC_HIP_CHECK(call)                                                                    \
    do {                                                                                      \
        const hipError_t err_ = (call);                                                       \
        if (err_ != hipSuccess) {                                                             \
            fprintf(stderr, "astaroth_amd: HIP error %s at %s:%d: %s\n", hipGetErrorName(err_), \
                    __FILE__, __LINE__, #call);                                               \
            return AC_FAILURE;                                                                \
        }                                                                                     \
    } while (0)

static AcResult fail(const char* msg)
{
    fprintf(stderr, "astaroth_amd: %s\n", msg);
    return AC_FAILURE;
}

namespace {
constexpr int kTimingSlots = 4096;
constexpr int kDiagSlots = DIAG_SLOTS;
}  // namespace


struct device_s {
    int id;
    AcMeshInfo local_config;
    hipStream_t streams[NUM_STREAMS];
    bool stream_created[NUM_STREAMS];
    // field buffers
    AcReal* in[NF];
    AcReal* out[NF];
    AcReal* w[NF];
    void* comp[NF];
    void* slab_in;
    void* slab_out;
    void* slab_w;
    void* slab_comp;
    int kahan_bits;
    size_t count;  // m^3
    // reductions and device-resident scalars
    double* partials;
    int max_partials;
    double* dres;                 // [0] reduce result, [1..] diagnostics
    unsigned long long* umax2;    // two slot pairs (x, y) of the timestep reductions
    double* time_acc;
    int cfl_cur;
    double cfl_courant;
    int cfl_mode;                 // 0 courant, 1 reference (2 when AC_additive_timestep)
    bool periodic_fused;          // substeps read the periodic images (no halo fill before them)
    bool forcing_time_acc;        // forcing start time = AC_current_time + time_acc - dt
    double* host_pinned;          // staging for small D2H reads
    // stencil coefficients (grid spacing folded in), set by acDeviceLoadStencilsFromConfig
    Coeffs<AcReal> C;
    // launch shape
    TiledShape shape;
    bool shape_tuned;
    // timing
    bool timing;
    int timing_used;
    hipEvent_t ev_start[kTimingSlots];
    hipEvent_t ev_stop[kTimingSlots];
    bool events_created;
};

static hipStream_t get_stream(Device d, Stream s)
{
    const int i = (s == STREAM_ALL || s < 0 || s >= NUM_STREAMS) ? 0 : (int)s;
    if (!d->stream_created[i]) {
        // Default priority only: streams of differing priority are not used anywhere.
        if (hipStreamCreateWithFlags(&d->streams[i], hipStreamNonBlocking) != hipSuccess)
            return nullptr;
        d->stream_created[i] = true;
    }
    return d->streams[i];
}

static bool volume_eq(const Volume a, const int3 b)
{
    return a.x == (size_t)b.x && a.y == (size_t)b.y && a.z == (size_t)b.z;
}

static void make_coeffs(const AcMeshInfo& info, Coeffs<AcReal>& C)
{
    const double c1[4] = {0.0, 3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0};
    const double c2[4] = {-49.0 / 18.0, 3.0 / 2.0, -3.0 / 20.0, 1.0 / 90.0};
    const double cu[4] = {-20.0 / 60.0, 15.0 / 60.0, -6.0 / 60.0, 1.0 / 60.0};
    const double cm[4] = {0.0, 270.0 / 720.0, -27.0 / 720.0, 2.0 / 720.0};
    const AcReal3 ds = info.real3_params[AC_ds];
    const double inv[3] = {1.0 / (double)ds.x, 1.0 / (double)ds.y, 1.0 / (double)ds.z};
    for (int a = 0; a < 3; ++a)
        for (int o = 0; o < 4; ++o) {
            C.d1[a][o] = (AcReal)(c1[o] * inv[a]);
            C.d2[a][o] = (AcReal)(c2[o] * inv[a] * inv[a]);
            C.up[a][o] = (AcReal)(cu[o] * inv[a]);
        }
    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int p = 0; p < 3; ++p)
        for (int o = 0; o < 4; ++o)
            C.mx[p][o] = (AcReal)(cm[o] * inv[pairs[p][0]] * inv[pairs[p][1]]);
}

static void make_phys(const AcMeshInfo& info, PhysParams<AcReal>& P)
{
    const double cs = info.real_params[AC_cs_sound];
    const double cp = info.real_params[AC_cp_sound];
    const double gamma = info.real_params[AC_gamma];
    const double mu0 = info.real_params[AC_mu0];
    P.cs2_sound = (AcReal)(cs * cs);
    P.nu = info.real_params[AC_nu_visc];
    P.zeta = info.real_params[AC_zeta];
    P.eta = info.real_params[AC_eta];
    P.mu0 = (AcReal)mu0;
    P.inv_mu0 = (AcReal)(1.0 / mu0);
    P.cp = (AcReal)cp;
    P.inv_cp = (AcReal)(1.0 / cp);
    P.gamma = (AcReal)gamma;
    P.gm1 = (AcReal)(gamma - 1.0);
    P.lnT0 = info.real_params[AC_lnT0];
    P.lnrho0 = info.real_params[AC_lnrho0];
    P.K = info.real_params[AC_K_heatcond];
    P.gamma_inv_cp = (AcReal)(gamma / cp);
    P.eta_mu0 = (AcReal)((double)info.real_params[AC_eta] * mu0);
    P.nu_third = (AcReal)((double)info.real_params[AC_nu_visc] / 3.0);
    P.pT_scale = (AcReal)(cs * cs * exp(-(double)info.real_params[AC_lnT0]));
}

// Parameters of the device timestep formula (cfl_dt) for the device's timestep mode. In the
// reference modes the Courant factors are AC_cdt and AC_cdtv of the configuration and the
// diffusive limit uses max(nu, eta) (default CFL: its chi term is 0, ref
// samples/standalone_mpi/main.cc:413) or max(nu, eta, gamma chi) (additive, ref mhdsolver.ac:560).
static void set_cfl_args(Device d, SolveArgs<AcReal>& a)
{
    const AcMeshInfo& info = d->local_config;
    const double nu = (double)info.real_params[AC_nu_visc], eta = (double)info.real_params[AC_eta];
    a.cfl_mode = d->cfl_mode;
    a.cfl_cs2 = (double)info.real_params[AC_cs_sound] * (double)info.real_params[AC_cs_sound];
    a.cfl_cdtv = (double)info.real_params[AC_cdtv];
    double visc = std::max(nu, eta);
    if (d->cfl_mode == 2)
        visc = std::max(visc, (double)info.real_params[AC_gamma] * (double)info.real_params[AC_chi]);
    a.cfl_visc = visc;
    if (d->cfl_mode != 0)
        a.cfl_cdt = (double)info.real_params[AC_cdt];
}

static void fill_solve_args(Device d, SolveArgs<AcReal>& a, int step, AcReal dt)
{
    const AcMeshInfo& info = d->local_config;
    const int3 n = info.int3_params[AC_nlocal];
    for (int q = 0; q < NF; ++q) {
        a.in[q] = d->in[q];
        a.out[q] = d->out[q];
        a.w[q] = d->w[q];
        a.comp[q] = d->comp[q];
    }
    a.nx = n.x;
    a.ny = n.y;
    a.nz = n.z;
    a.mx = n.x + 2 * R;
    a.my = n.y + 2 * R;
    a.mz = n.z + 2 * R;
    a.sz = (long long)a.mx * a.my;
    a.fstride = (long long)d->count;
    a.C = d->C;
    make_phys(info, a.P);
    const AcReal alpha[3] = {(AcReal)0.0, (AcReal)(-5.0 / 9.0), (AcReal)(-153.0 / 128.0)};
    const AcReal beta[3] = {(AcReal)(1.0 / 3.0), (AcReal)(15.0 / 16.0), (AcReal)(8.0 / 15.0)};
    a.dt = dt;
    a.alpha = alpha[step];
    a.beta = beta[step];
    a.step = step;
    a.umax2_cur = nullptr;
    a.umax2_next = nullptr;
    a.time_acc = nullptr;
    a.cfl_cdt = 0.0;
    a.cfl_dsmin = (double)info.real_params[AC_dsmin];
    a.cfl_cs = (double)info.real_params[AC_cs_sound];
    set_cfl_args(d, a);
    a.wrap = d->periodic_fused ? 1 : 0;
}

// Launch shapes: kind 0 runs the delay-line (or ring) kernel of solve_tiled.hip, kind 1 the
// unrolled sweep of solve_sweep.hip; tx == 0 is the direct kernel.
static bool shape_fits(TiledShape s, int3 n)
{
    return s.kind == 1 ? sweep_shape_fits(s, n.x, n.y, n.z) : tiled_shape_fits(s, n.x, n.y, n.z);
}

// Whether the kernel that runs for shape s reads the periodic images itself (honours a.wrap).
static bool shape_reads_periodic(const SolveArgs<AcReal>& a, int kb, TiledShape s)
{
    if (s.tx <= 0)
        return false;
    return s.kind == 1 ? sweep_applies<AcReal>(a, kb, s) : tiled_reads_periodic<AcReal>(a, kb, s);
}

static bool launch_shape(const SolveArgs<AcReal>& a, bool upwd, bool entropy, int kb, TiledShape s,
                         hipStream_t st)
{
    if (s.tx <= 0)
        return false;
    return s.kind == 1 ? launch_solve_sweep<AcReal>(a, upwd, entropy, kb, s, st)
                       : launch_solve_tiled<AcReal>(a, upwd, entropy, kb, s, st);
}

// The shape the configuration pins (AC_solve_tile_x > 0), or the direct kernel's (tx < 0).
static TiledShape config_shape(const AcMeshInfo& c)
{
    return TiledShape{c.int_params[AC_solve_tile_x], c.int_params[AC_solve_tile_y],
                      c.int_params[AC_solve_zchunk], c.int_params[AC_solve_kernel] == 1 ? 1 : 0};
}

static AcResult launch_solve(Device d, const SolveArgs<AcReal>& a_in, hipStream_t s)
{
    const AcMeshInfo& info = d->local_config;
    const bool upwd = info.bool_params[AC_lupwd] != 0;
    const bool entropy = info.bool_params[AC_lentropy] != 0;
    const int kb = d->kahan_bits;
    SolveArgs<AcReal> a = a_in;
    if (a.wrap && !shape_reads_periodic(a, kb, d->shape)) {
        // the kernel that runs reads the halos: fill them first
        const int3 n = info.int3_params[AC_nlocal];
        launch_periodic_fill<AcReal>(d->in, NF, n.x, n.y, n.z, s);
        a.wrap = 0;
    }
    const bool timed = d->timing && d->events_created && d->timing_used < kTimingSlots;
    if (timed)
        AC_HIP_CHECK(hipEventRecord(d->ev_start[d->timing_used], s));
    if (!launch_shape(a, upwd, entropy, kb, d->shape, s))
        launch_solve_direct<AcReal>(a, upwd, entropy, kb, s);
    AC_HIP_CHECK(hipGetLastError());
    if (timed) {
        AC_HIP_CHECK(hipEventRecord(d->ev_stop[d->timing_used], s));
        ++d->timing_used;
    }
    return AC_SUCCESS;
}

// The helical forcing of substep 2 (ref mhdsolver.ac:614-631): added to u in `out` after the
// solve. With umax2_next set, the forcing kernel reduces max |u|^2 of the forced state instead
// of the solve.
static void launch_forcing_step(Device d, const SolveArgs<AcReal>& a, hipStream_t s,
                                unsigned long long* umax2_next)
{
    const AcMeshInfo& info = d->local_config;
    const AcReal* r = info.real_params;
    ForcingArgs<AcReal> f;
    memset((void*)&f, 0, sizeof(f));
    for (int q = 0; q < 3; ++q)
        f.uu[q] = d->out[VTXBUF_UUX + q];
    f.nx = a.nx;
    f.ny = a.ny;
    f.nz = a.nz;
    f.mx = a.mx;
    f.my = a.my;
    f.sz = a.sz;
    const int3 off = info.int3_params[AC_multigpu_offset];
    const int3 ng = info.int3_params[AC_ngrid];
    f.gx0 = off.x;
    f.gy0 = off.y;
    f.gz0 = off.z;
    f.ngx = ng.x;
    f.ngy = ng.y;
    f.ngz = ng.z;
    const AcReal3 ds = info.real3_params[AC_ds];
    f.ds[0] = (double)ds.x;
    f.ds[1] = (double)ds.y;
    f.ds[2] = (double)ds.z;
    f.k[0] = (double)r[AC_k_forcex];
    f.k[1] = (double)r[AC_k_forcey];
    f.k[2] = (double)r[AC_k_forcez];
    f.fre[0] = (double)r[AC_ff_hel_rex];
    f.fre[1] = (double)r[AC_ff_hel_rey];
    f.fre[2] = (double)r[AC_ff_hel_rez];
    f.fim[0] = (double)r[AC_ff_hel_imx];
    f.fim[1] = (double)r[AC_ff_hel_imy];
    f.fim[2] = (double)r[AC_ff_hel_imz];
    f.phase = (double)r[AC_forcing_phase];
    const double cs = sqrt((double)r[AC_cs2_sound]);
    f.nn = cs * (double)r[AC_forcing_magnitude] * sqrt((double)r[AC_kaver] * cs);
    f.t_base = (double)r[AC_current_time];
    f.switch_time = (double)r[AC_switch_forcing];
    f.time_acc = d->forcing_time_acc ? a.time_acc : nullptr;
    f.s = a;
    f.umax2_next = umax2_next;
    launch_forcing<AcReal>(f, s);
}

static bool forcing_on(Device d) { return d->local_config.bool_params[AC_lforcing] != 0; }

extern "C" {

AcResult acDeviceCreate(const int id, const AcMeshInfo device_config, Device* device_handle)
{
    *device_handle = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail("acDeviceCreate: no HIP device available");
    if (id < 0 || id >= count)
        return fail("acDeviceCreate: device id out of range");
    AC_HIP_CHECK(hipSetDevice(id));

    device_s* d = new device_s();
    memset((void*)d, 0, sizeof(*d));
    d->id = id;
    d->local_config = device_config;
    int3& off = d->local_config.int3_params[AC_multigpu_offset];
    if (off.x < 0 || off.y < 0 || off.z < 0) {
        fprintf(stderr, "astaroth_amd: invalid AC_multigpu_offset, using (0,0,0)\n");
        off.x = off.y = off.z = 0;
    }
    if (acHostUpdateParams(&d->local_config) != AC_SUCCESS) {
        delete d;
        return fail("acDeviceCreate: invalid grid configuration (AC_ngrid / AC_ds)");
    }
    const int3 m = d->local_config.int3_params[AC_mlocal];
    d->count = (size_t)m.x * (size_t)m.y * (size_t)m.z;
    int kb = d->local_config.int_params[AC_kahan_bits];
    if (sizeof(AcReal) == 8)
        kb = 0;
    if (kb != 0 && kb != 8 && kb != 16 && kb != 32) {
        delete d;
        return fail("acDeviceCreate: AC_kahan_bits must be 0, 8, 16 or 32");
    }
    d->kahan_bits = kb;
    make_coeffs(d->local_config, d->C);

    hipStream_t s0 = get_stream(d, STREAM_0);
    if (!s0) {
        delete d;
        return fail("acDeviceCreate: stream creation failed");
    }
    const size_t bytes = d->count * sizeof(AcReal) * NF;
    if (hipMalloc(&d->slab_in, bytes) != hipSuccess || hipMalloc(&d->slab_out, bytes) != hipSuccess ||
        hipMalloc(&d->slab_w, bytes) != hipSuccess) {
        hipFree(d->slab_in);
        hipFree(d->slab_out);
        hipFree(d->slab_w);
        delete d;
        return fail("acDeviceCreate: device allocation failed");
    }
    size_t comp_bytes = 0;
    if (kb > 0) {
        comp_bytes = d->count * (size_t)(kb / 8) * NF;
        if (hipMalloc(&d->slab_comp, comp_bytes) != hipSuccess) {
            hipFree(d->slab_in);
            hipFree(d->slab_out);
            hipFree(d->slab_w);
            delete d;
            return fail("acDeviceCreate: device allocation of the compensation slab failed");
        }
    }
    for (int q = 0; q < NF; ++q) {
        d->in[q] = (AcReal*)d->slab_in + (size_t)q * d->count;
        d->out[q] = (AcReal*)d->slab_out + (size_t)q * d->count;
        d->w[q] = (AcReal*)d->slab_w + (size_t)q * d->count;
        // one record of NF compensations per cell: field q's are every NF-th, from byte q kb / 8
        d->comp[q] = kb > 0 ? (void*)((char*)d->slab_comp + (size_t)q * (kb / 8)) : nullptr;
    }
    launch_zero_bytes(d->slab_in, bytes, s0);
    launch_zero_bytes(d->slab_out, bytes, s0);
    launch_zero_bytes(d->slab_w, bytes, s0);
    if (kb > 0)
        launch_zero_bytes(d->slab_comp, comp_bytes, s0);

    const int3 n = d->local_config.int3_params[AC_nlocal];
    d->max_partials = std::max(reduce_num_partials(n.x, n.y, n.z), 1) * kDiagSlots;
    AC_HIP_CHECK(hipMalloc(&d->partials, sizeof(double) * d->max_partials));
    AC_HIP_CHECK(hipMalloc(&d->dres, sizeof(double) * (1 + kDiagSlots)));
    AC_HIP_CHECK(hipMalloc(&d->umax2, sizeof(unsigned long long) * 4));
    AC_HIP_CHECK(hipMalloc(&d->time_acc, sizeof(double)));
    launch_zero_bytes(d->umax2, sizeof(unsigned long long) * 4, s0);
    launch_zero_bytes(d->time_acc, sizeof(double), s0);
    AC_HIP_CHECK(hipHostMalloc((void**)&d->host_pinned, sizeof(double) * (2 + kDiagSlots)));

    // AC_solve_tile_x > 0 pins a tiled shape (AC_solve_kernel selects the kernel), < 0 pins the
    // direct kernel, 0 autotunes
    d->shape = config_shape(d->local_config);
    d->shape_tuned = d->shape.tx != 0;
    if (d->shape.tx < 0 || (d->shape.tx > 0 && !shape_fits(d->shape, n)))
        d->shape = TiledShape{0, 0, 0, 0};
    AC_HIP_CHECK(hipGetLastError());
    AC_HIP_CHECK(hipStreamSynchronize(s0));
    *device_handle = d;
    return AC_SUCCESS;
}

AcResult acDeviceDestroy(Device* device_ptr)
{
    Device d = *device_ptr;
    if (!d)
        return AC_SUCCESS;
    hipSetDevice(d->id);
    hipDeviceSynchronize();
    hipFree(d->slab_in);
    hipFree(d->slab_out);
    hipFree(d->slab_w);
    if (d->slab_comp)
        hipFree(d->slab_comp);
    hipFree(d->partials);
    hipFree(d->dres);
    hipFree(d->umax2);
    hipFree(d->time_acc);
    hipHostFree(d->host_pinned);
    if (d->events_created)
        for (int i = 0; i < kTimingSlots; ++i) {
            hipEventDestroy(d->ev_start[i]);
            hipEventDestroy(d->ev_stop[i]);
        }
    for (int i = 0; i < NUM_STREAMS; ++i)
        if (d->stream_created[i])
            hipStreamDestroy(d->streams[i]);
    delete d;
    *device_ptr = nullptr;
    return AC_SUCCESS;
}

AcResult acDeviceResetMesh(const Device d, const Stream stream)
{
    hipSetDevice(d->id);
    hipStream_t s = get_stream(d, stream);
    const size_t bytes = d->count * sizeof(AcReal) * NF;
    launch_zero_bytes(d->slab_in, bytes, s);
    launch_zero_bytes(d->slab_out, bytes, s);
    launch_zero_bytes(d->slab_w, bytes, s);
    if (d->slab_comp)
        launch_zero_bytes(d->slab_comp, d->count * (size_t)(d->kahan_bits / 8) * NF, s);
    AC_HIP_CHECK(hipGetLastError());
    return AC_SUCCESS;
}

AcResult acDevicePrintInfo(const Device d)
{
    hipDeviceProp_t p;
    AC_HIP_CHECK(hipGetDeviceProperties(&p, d->id));
    const int3 n = d->local_config.int3_params[AC_nlocal];
    printf("Device %d: %s (%s), %d CUs, %.1f GiB\n", d->id, p.name, p.gcnArchName,
           p.multiProcessorCount, p.totalGlobalMem / 1073741824.0);
    printf("  grid %d x %d x %d, AcReal %zu bytes, kahan bits %d, solve shape %d x %d x %d (%s)\n",
           n.x, n.y, n.z, sizeof(AcReal), d->kahan_bits, d->shape.tx, d->shape.ty, d->shape.zchunk,
           d->shape.tx <= 0 ? "direct" : (d->shape.kind == 1 ? "sweep" : "delay-line"));
    return AC_SUCCESS;
}

AcResult acDeviceSynchronizeStream(const Device d, const Stream stream)
{
    AC_HIP_CHECK(hipSetDevice(d->id));
    if (stream == STREAM_ALL) {
        AC_HIP_CHECK(hipDeviceSynchronize());
    }
    else {
        AC_HIP_CHECK(hipStreamSynchronize(get_stream(d, stream)));
    }
    return AC_SUCCESS;
}

AcResult acDeviceSwapBuffer(const Device d, const VertexBufferHandle h)
{
    if ((int)h < 0 || (int)h >= NF)
        return fail("acDeviceSwapBuffer: invalid handle");
    std::swap(d->in[h], d->out[h]);
    return AC_SUCCESS;
}

AcResult acDeviceSwapBuffers(const Device d)
{
    for (int q = 0; q < NF; ++q)
        std::swap(d->in[q], d->out[q]);
    return AC_SUCCESS;
}

AcResult acDeviceLoadScalarUniform(const Device d, const Stream, const AcRealParam param,
                                   const AcReal value)
{
    if ((int)param < 0 || (int)param >= NUM_REAL_PARAMS)
        return fail("acDeviceLoadScalarUniform: invalid parameter");
    d->local_config.real_params[param] = value;
    return acHostUpdateParams(&d->local_config);
}

AcResult acDeviceLoadForcingParams(const Device d, const Stream, const ForcingParams fp)
{
    loadForcingParamsToMeshInfo(fp, &d->local_config);
    return AC_SUCCESS;
}

AcResult acDeviceLoadVectorUniform(const Device d, const Stream, const AcReal3Param param,
                                   const AcReal3 value)
{
    if ((int)param < 0 || (int)param >= NUM_REAL3_PARAMS)
        return fail("acDeviceLoadVectorUniform: invalid parameter");
    d->local_config.real3_params[param] = value;
    return acHostUpdateParams(&d->local_config);
}

AcResult acDeviceLoadIntUniform(const Device d, const Stream, const AcIntParam param, const int value)
{
    if ((int)param < 0 || (int)param >= NUM_INT_PARAMS)
        return fail("acDeviceLoadIntUniform: invalid parameter");
    if (param == AC_kahan_bits && value != d->kahan_bits)
        return fail("acDeviceLoadIntUniform: AC_kahan_bits is fixed at acDeviceCreate");
    d->local_config.int_params[param] = value;
    if (param == AC_solve_tile_x || param == AC_solve_tile_y || param == AC_solve_zchunk ||
        param == AC_solve_kernel) {
        const int3 n = d->local_config.int3_params[AC_nlocal];
        const TiledShape s = config_shape(d->local_config);
        if (s.tx > 0 && shape_fits(s, n)) {
            d->shape = s;
            d->shape_tuned = true;
        }
        else if (s.tx <= 0) {
            d->shape = TiledShape{0, 0, 0, 0};
            d->shape_tuned = s.tx < 0;
        }
        else {
            return fail("acDeviceLoadIntUniform: solve launch shape does not fit the grid");
        }
    }
    return AC_SUCCESS;
}

AcResult acDeviceLoadBoolUniform(const Device d, const Stream, const AcBoolParam param,
                                 const bool value)
{
    if ((int)param < 0 || (int)param >= NUM_BOOL_PARAMS)
        return fail("acDeviceLoadBoolUniform: invalid parameter");
    d->local_config.bool_params[param] = value ? 1 : 0;
    return AC_SUCCESS;
}

AcResult acDeviceLoadInt3Uniform(const Device d, const Stream, const AcInt3Param param,
                                 const int3 value)
{
    if ((int)param < 0 || (int)param >= NUM_INT3_PARAMS)
        return fail("acDeviceLoadInt3Uniform: invalid parameter");
    if (param == AC_ngrid || param == AC_nlocal || param == AC_mlocal || param == AC_mgrid ||
        param == AC_nmin || param == AC_nlocal_max || param == AC_ngrid_max)
        return fail("acDeviceLoadInt3Uniform: grid dimensions are fixed at acDeviceCreate");
    d->local_config.int3_params[param] = value;
    return AC_SUCCESS;
}

AcResult acDeviceStoreScalarUniform(const Device d, const Stream, const AcRealParam param,
                                    AcReal* value)
{
    if ((int)param < 0 || (int)param >= NUM_REAL_PARAMS)
        return fail("acDeviceStoreScalarUniform: invalid parameter");
    *value = d->local_config.real_params[param];
    return AC_SUCCESS;
}

AcResult acDeviceStoreVectorUniform(const Device d, const Stream, const AcReal3Param param,
                                    AcReal3* value)
{
    if ((int)param < 0 || (int)param >= NUM_REAL3_PARAMS)
        return fail("acDeviceStoreVectorUniform: invalid parameter");
    *value = d->local_config.real3_params[param];
    return AC_SUCCESS;
}

AcResult acDeviceLoadMeshInfo(const Device d, const AcMeshInfo config)
{
    AcMeshInfo c = config;
    if (acHostUpdateParams(&c) != AC_SUCCESS)
        return fail("acDeviceLoadMeshInfo: invalid configuration");
    for (int p = AC_ngrid; p <= AC_ngrid_max; ++p) {
        const int3 a = c.int3_params[p], b = d->local_config.int3_params[p];
        if (a.x != b.x || a.y != b.y || a.z != b.z)
            return fail("acDeviceLoadMeshInfo: grid dimensions differ from acDeviceCreate");
    }
    c.int_params[AC_kahan_bits] = d->kahan_bits;
    d->local_config = c;
    return AC_SUCCESS;
}

// ref include/astaroth_device.h:229, src/core/device.cc:447-473: evaluates the stencil
// coefficients from the device's configuration (AC_ds) and makes them the ones every later
// substep, reduction and diagnostic uses. acDeviceCreate does this once; a caller that changes
// the grid spacing afterwards (acDeviceLoadVectorUniform, acDeviceLoadMeshInfo) calls it again,
// as with the reference, whose kernels keep the coefficients last uploaded.
AcResult acDeviceLoadStencilsFromConfig(const Device d, const Stream)
{
    const AcReal3 ds = d->local_config.real3_params[AC_ds];
    if (!(std::isfinite((double)ds.x) && std::isfinite((double)ds.y) && std::isfinite((double)ds.z) &&
          ds.x > 0 && ds.y > 0 && ds.z > 0))
        return fail("acDeviceLoadStencilsFromConfig: AC_ds must be finite and positive");
    make_coeffs(d->local_config, d->C);
    return AC_SUCCESS;
}

AcMeshInfo acDeviceGetConfig(const Device d) { return d->local_config; }
AcMeshInfo acDeviceGetLocalConfig(const Device d) { return d->local_config; }

static size_t host_index(const int3 p, const AcMeshInfo& info)
{
    const int3 m = info.int3_params[AC_mlocal];
    return (size_t)p.x + (size_t)m.x * ((size_t)p.y + (size_t)m.y * (size_t)p.z);
}

static size_t info_count(const AcMeshInfo& info)
{
    const int3 m = info.int3_params[AC_mlocal];
    return (size_t)m.x * (size_t)m.y * (size_t)m.z;
}

AcResult acDeviceLoadVertexBufferWithOffset(const Device d, const Stream stream,
                                            const AcMesh host_mesh, const VertexBufferHandle h,
                                            const int3 src, const int3 dst, const int num_vertices)
{
    if ((int)h < 0 || (int)h >= NF)
        return fail("acDeviceLoadVertexBufferWithOffset: invalid handle");
    if (!host_mesh.vertex_buffer[h])
        return AC_NOT_ALLOCATED;
    AC_HIP_CHECK(hipSetDevice(d->id));
    const size_t si = host_index(src, host_mesh.info);
    const size_t di = host_index(dst, d->local_config);
    if (num_vertices < 0 || di + (size_t)num_vertices > d->count ||
        si + (size_t)num_vertices > info_count(host_mesh.info))
        return fail("acDeviceLoadVertexBufferWithOffset: range outside the mesh");
    hipStream_t s = get_stream(d, stream);
    AC_HIP_CHECK(hipMemcpyAsync(d->in[h] + di, host_mesh.vertex_buffer[h] + si,
                                sizeof(AcReal) * (size_t)num_vertices, hipMemcpyHostToDevice, s));
    if (d->kahan_bits > 0)
        launch_zero_comp_field(d->slab_comp, d->count, d->kahan_bits, (int)h, s);
    return AC_SUCCESS;
}

AcResult acDeviceLoadMeshWithOffset(const Device d, const Stream stream, const AcMesh host_mesh,
                                    const int3 src, const int3 dst, const int num_vertices)
{
    int r = 0;
    for (int q = 0; q < NF; ++q)
        r |= acDeviceLoadVertexBufferWithOffset(d, stream, host_mesh, (VertexBufferHandle)q, src,
                                                dst, num_vertices);
    return (AcResult)r;
}

AcResult acDeviceLoadVertexBuffer(const Device d, const Stream stream, const AcMesh host_mesh,
                                  const VertexBufferHandle h)
{
    if (info_count(host_mesh.info) != d->count)
        return fail("acDeviceLoadVertexBuffer: host and device mesh sizes differ");
    int3 z;
    z.x = z.y = z.z = 0;
    return acDeviceLoadVertexBufferWithOffset(d, stream, host_mesh, h, z, z, (int)d->count);
}

AcResult acDeviceLoadMesh(const Device d, const Stream stream, const AcMesh host_mesh)
{
    int r = 0;
    for (int q = 0; q < NF; ++q)
        r |= acDeviceLoadVertexBuffer(d, stream, host_mesh, (VertexBufferHandle)q);
    return (AcResult)r;
}

extern "C++" {
namespace acamd {
void fourier_tables(const AcMeshInfo& info, uint64_t seed, int nmodes, int kmax, double ampl,
                    std::vector<double>& head, std::vector<double>& tab);
uint64_t lnrho_seed(uint64_t seed);
}  // namespace acamd
}

AcResult acDeviceInitFourier(const Device d, const Stream stream, const unsigned long long seed,
                             const int nmodes, const int kmax, const double ampl,
                             const double lnrho_perturb)
{
    if (nmodes < 0 || kmax < 1)
        return fail("acDeviceInitFourier: nmodes >= 0 and kmax >= 1 required");
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    const AcMeshInfo& info = d->local_config;
    std::vector<double> head, tab;
    fourier_tables(info, seed, nmodes, kmax, ampl, head, tab);
    double *dhead = nullptr, *dtab = nullptr;
    const size_t hb = std::max<size_t>(head.size(), 1) * sizeof(double);
    const size_t tb = std::max<size_t>(tab.size(), 1) * sizeof(double);
    if (hipMalloc(&dhead, hb) != hipSuccess || hipMalloc(&dtab, tb) != hipSuccess) {
        hipFree(dhead);
        return fail("acDeviceInitFourier: device allocation of the mode tables failed");
    }
    AC_HIP_CHECK(hipMemcpyAsync(dhead, head.data(), head.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
    AC_HIP_CHECK(hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
    const int3 n = info.int3_params[AC_nlocal];
    FourierArgs<AcReal> A;
    for (int q = 0; q < 3; ++q) {
        A.uu[q] = d->in[VTXBUF_UUX + q];
        A.aa[q] = d->in[VTXBUF_AX + q];
    }
    A.lnrho = d->in[VTXBUF_LNRHO];
    A.ss = d->in[VTXBUF_ENTROPY];
    A.head = dhead;
    A.tab = dtab;
    A.nmodes = nmodes;
    A.nx = n.x;
    A.ny = n.y;
    A.nz = n.z;
    A.mx = n.x + 2 * R;
    A.my = n.y + 2 * R;
    A.mz = n.z + 2 * R;
    A.lnrho_perturb = lnrho_perturb;
    A.lseed = lnrho_seed(seed);
    launch_fourier_ic<AcReal>(A, s);
    AcReal* f[NF];
    for (int q = 0; q < NF; ++q)
        f[q] = d->in[q];
    launch_periodic_fill<AcReal>(f, NF, n.x, n.y, n.z, s);
    for (int q = 0; q < NF; ++q) {
        launch_fill<AcReal>(d->out[q], d->count, (AcReal)0, s);
        if (d->kahan_bits > 0)
            launch_zero_comp_field(d->slab_comp, d->count, d->kahan_bits, q, s);
    }
    AC_HIP_CHECK(hipGetLastError());
    // the tables are consumed by the kernel above; free them once it has run
    AC_HIP_CHECK(hipStreamSynchronize(s));
    hipFree(dhead);
    hipFree(dtab);
    return AC_SUCCESS;
}

AcResult acDeviceSetVertexBuffer(const Device d, const Stream stream, const VertexBufferHandle h,
                                 const AcReal value)
{
    if ((int)h < 0 || (int)h >= NF)
        return fail("acDeviceSetVertexBuffer: invalid handle");
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    launch_fill<AcReal>(d->in[h], d->count, value, s);
    launch_fill<AcReal>(d->out[h], d->count, value, s);
    if (d->kahan_bits > 0)
        launch_zero_comp_field(d->slab_comp, d->count, d->kahan_bits, (int)h, s);
    AC_HIP_CHECK(hipGetLastError());
    return AC_SUCCESS;
}

AcResult acDeviceFlushOutputBuffers(const Device d, const Stream stream)
{
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    for (int q = 0; q < NF; ++q)
        launch_fill<AcReal>(d->out[q], d->count, (AcReal)0, s);
    AC_HIP_CHECK(hipGetLastError());
    return AC_SUCCESS;
}

AcResult acDeviceStoreVertexBufferWithOffset(const Device d, const Stream stream,
                                             const VertexBufferHandle h, const int3 src,
                                             const int3 dst, const int num_vertices,
                                             AcMesh* host_mesh)
{
    if ((int)h < 0 || (int)h >= NF)
        return fail("acDeviceStoreVertexBufferWithOffset: invalid handle");
    if (!host_mesh->vertex_buffer[h])
        return AC_NOT_ALLOCATED;
    AC_HIP_CHECK(hipSetDevice(d->id));
    const size_t si = host_index(src, d->local_config);
    const size_t di = host_index(dst, host_mesh->info);
    if (num_vertices < 0 || si + (size_t)num_vertices > d->count ||
        di + (size_t)num_vertices > info_count(host_mesh->info))
        return fail("acDeviceStoreVertexBufferWithOffset: range outside the mesh");
    hipStream_t s = get_stream(d, stream);
    AC_HIP_CHECK(hipMemcpyAsync(host_mesh->vertex_buffer[h] + di, d->in[h] + si,
                                sizeof(AcReal) * (size_t)num_vertices, hipMemcpyDeviceToHost, s));
    return AC_SUCCESS;
}

AcResult acDeviceStoreMeshWithOffset(const Device d, const Stream stream, const int3 src,
                                     const int3 dst, const int num_vertices, AcMesh* host_mesh)
{
    int r = 0;
    for (int q = 0; q < NF; ++q)
        r |= acDeviceStoreVertexBufferWithOffset(d, stream, (VertexBufferHandle)q, src, dst,
                                                 num_vertices, host_mesh);
    return (AcResult)r;
}

AcResult acDeviceStoreVertexBuffer(const Device d, const Stream stream, const VertexBufferHandle h,
                                   AcMesh* host_mesh)
{
    if (info_count(host_mesh->info) != d->count)
        return fail("acDeviceStoreVertexBuffer: host and device mesh sizes differ");
    int3 z;
    z.x = z.y = z.z = 0;
    return acDeviceStoreVertexBufferWithOffset(d, stream, h, z, z, (int)d->count, host_mesh);
}

AcResult acDeviceStoreMesh(const Device d, const Stream stream, AcMesh* host_mesh)
{
    int r = 0;
    for (int q = 0; q < NF; ++q)
        r |= acDeviceStoreVertexBuffer(d, stream, (VertexBufferHandle)q, host_mesh);
    return (AcResult)r;
}

static AcResult check_interior(Device d, const Volume start, const Volume end, const char* who)
{
    const AcMeshInfo& info = d->local_config;
    if (!volume_eq(start, info.int3_params[AC_nmin]) || !volume_eq(end, info.int3_params[AC_nlocal_max])) {
        fprintf(stderr, "astaroth_amd: %s: start/end must be the full interior [nmin, nlocal_max)\n", who);
        return AC_FAILURE;
    }
    return AC_SUCCESS;
}

static AcResult ensure_shape(Device d, hipStream_t s, bool w_dead);

AcResult acDeviceIntegrateSubstep(const Device d, const Stream stream, const int step_number,
                                  const Volume start, const Volume end, const AcReal dt)
{
    if (check_interior(d, start, end, "acDeviceIntegrateSubstep") != AC_SUCCESS)
        return AC_FAILURE;
    if (step_number < 0 || step_number > 2)
        return fail("acDeviceIntegrateSubstep: step_number must be 0, 1 or 2");
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    if (ensure_shape(d, s, step_number == 0) != AC_SUCCESS)
        return AC_FAILURE;
    SolveArgs<AcReal> a;
    fill_solve_args(d, a, step_number, dt);
    const AcResult r = launch_solve(d, a, s);
    if (r == AC_SUCCESS && step_number == 2 && forcing_on(d))
        launch_forcing_step(d, a, s, nullptr);
    return r;
}

// The box [start, end) of the m^3 arrays as int bounds; false if it leaves [0, mlocal).
static bool fill_box(Device d, const Volume start, const Volume end, int lo[3], int hi[3])
{
    const int3 m = d->local_config.int3_params[AC_mlocal];
    const size_t s3[3] = {start.x, start.y, start.z}, e3[3] = {end.x, end.y, end.z};
    const int m3[3] = {m.x, m.y, m.z};
    for (int a = 0; a < 3; ++a) {
        if (s3[a] > e3[a] || e3[a] > (size_t)m3[a])
            return false;
        lo[a] = (int)s3[a];
        hi[a] = (int)e3[a];
    }
    return true;
}

// Fills the halo cells inside [start, end) of the given fields: the full volume [0, mlocal) in
// the one-launch fill over halo cells only, any other box in the box kernel (ref
// src/core/device.cc:998-1017, acc-runtime/built-in/kernels.h:17-41).
static AcResult periodic_fill(Device d, const Stream stream, AcReal* const* f, int nf,
                              const Volume start, const Volume end, const char* who)
{
    int lo[3], hi[3];
    if (!fill_box(d, start, end, lo, hi)) {
        fprintf(stderr, "astaroth_amd: %s: start/end must satisfy start <= end <= mlocal\n", who);
        return AC_FAILURE;
    }
    AC_HIP_CHECK(hipSetDevice(d->id));
    const int3 n = d->local_config.int3_params[AC_nlocal];
    const int3 m = d->local_config.int3_params[AC_mlocal];
    hipStream_t s = get_stream(d, stream);
    if (lo[0] == 0 && lo[1] == 0 && lo[2] == 0 && hi[0] == m.x && hi[1] == m.y && hi[2] == m.z)
        launch_periodic_fill<AcReal>(f, nf, n.x, n.y, n.z, s);
    else
        launch_periodic_fill_box<AcReal>(f, nf, n.x, n.y, n.z, lo, hi, s);
    AC_HIP_CHECK(hipGetLastError());
    return AC_SUCCESS;
}

AcResult acDevicePeriodicBoundcondStep(const Device d, const Stream stream,
                                       const VertexBufferHandle h, const Volume start,
                                       const Volume end)
{
    if ((int)h < 0 || (int)h >= NF)
        return fail("acDevicePeriodicBoundcondStep: invalid handle");
    AcReal* f[1] = {d->in[h]};
    return periodic_fill(d, stream, f, 1, start, end, "acDevicePeriodicBoundcondStep");
}

AcResult acDevicePeriodicBoundconds(const Device d, const Stream stream, const Volume start,
                                    const Volume end)
{
    return periodic_fill(d, stream, d->in, NF, start, end, "acDevicePeriodicBoundconds");
}

// The two-stage reduction with its partial results in `partials` (max_partials of them) and the
// final value in *result_dev, both device memory.
static AcResult reduce_into(Device d, const Stream stream, const AcReduction r, int map,
                            const VertexBufferHandle* h, int nh, bool post, double* partials,
                            int max_partials, double* result_dev, AcReal* result)
{
    for (int q = 0; q < nh; ++q)
        if ((int)h[q] < 0 || (int)h[q] >= NF)
            return fail("acDeviceReduce: invalid handle");
    if (map == AC_NULL_KERNEL || r.reduce_op == AC_REDUCE_OP_NO_REDUCE)
        return fail("acDeviceReduce: reduction type not defined for this arity");
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    const AcReal* bufs[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int q = 0; q < nh; ++q)
        bufs[q] = d->in[h[q]];
    const int3 n = d->local_config.int3_params[AC_nlocal];
    launch_reduce<AcReal>(bufs, (int)map, (int)r.reduce_op, n.x, n.y, n.z, partials, max_partials,
                          result_dev, s);
    AC_HIP_CHECK(hipGetLastError());
    AC_HIP_CHECK(hipMemcpyAsync(d->host_pinned, result_dev, sizeof(double), hipMemcpyDeviceToHost, s));
    AC_HIP_CHECK(hipStreamSynchronize(s));
    double v = d->host_pinned[0];
    if (post && r.post_processing_op == AC_RMS)
        v = sqrt(v / ((double)n.x * n.y * n.z));
    *result = (AcReal)v;
    return AC_SUCCESS;
}

static AcResult reduce_common(Device d, const Stream stream, const AcReduction r, int map,
                              const VertexBufferHandle* h, int nh, bool post, AcReal* result)
{
    return reduce_into(d, stream, r, map, h, nh, post, d->partials, d->max_partials, d->dres, result);
}

AcResult acDeviceReduceScalNoPostProcessing(const Device d, const Stream stream,
                                            const AcReduction r, const VertexBufferHandle h,
                                            AcReal* result)
{
    return reduce_common(d, stream, r, r.map_vtxbuf_single, &h, 1, false, result);
}

AcResult acDeviceReduceScal(const Device d, const Stream stream, const AcReduction r,
                            const VertexBufferHandle h, AcReal* result)
{
    return reduce_common(d, stream, r, r.map_vtxbuf_single, &h, 1, true, result);
}

AcResult acDeviceReduceVecNoPostProcessing(const Device d, const Stream stream, const AcReduction r,
                                           const VertexBufferHandle a, const VertexBufferHandle b,
                                           const VertexBufferHandle c, AcReal* result)
{
    const VertexBufferHandle h[3] = {a, b, c};
    return reduce_common(d, stream, r, r.map_vtxbuf_vec, h, 3, false, result);
}

AcResult acDeviceReduceVec(const Device d, const Stream stream, const AcReduction r,
                           const VertexBufferHandle a, const VertexBufferHandle b,
                           const VertexBufferHandle c, AcReal* result)
{
    const VertexBufferHandle h[3] = {a, b, c};
    return reduce_common(d, stream, r, r.map_vtxbuf_vec, h, 3, true, result);
}

AcResult acDeviceReduceVecScalNoPostProcessing(const Device d, const Stream stream,
                                               const AcReduction r, const VertexBufferHandle a,
                                               const VertexBufferHandle b,
                                               const VertexBufferHandle c,
                                               const VertexBufferHandle e, AcReal* result)
{
    const VertexBufferHandle h[4] = {a, b, c, e};
    return reduce_common(d, stream, r, r.map_vtxbuf_vec_scal, h, 4, false, result);
}

AcResult acDeviceReduceVecScal(const Device d, const Stream stream, const AcReduction r,
                               const VertexBufferHandle a, const VertexBufferHandle b,
                               const VertexBufferHandle c, const VertexBufferHandle e,
                               AcReal* result)
{
    const VertexBufferHandle h[4] = {a, b, c, e};
    return reduce_common(d, stream, r, r.map_vtxbuf_vec_scal, h, 4, true, result);
}

// ------------------------------------------------------------------ extensions

AcResult acDeviceSetTimestepMode(const Device d, const int mode)
{
    if (mode != AC_TIMESTEP_COURANT && mode != AC_TIMESTEP_REFERENCE)
        return fail("acDeviceSetTimestepMode: unknown mode");
    d->cfl_mode = mode == AC_TIMESTEP_COURANT
                      ? 0
                      : (d->local_config.bool_params[AC_additive_timestep] ? 2 : 1);
    return AC_SUCCESS;
}

AcResult acDeviceUpdateTimestepCFL(const Device d, const Stream stream, const AcReal courant)
{
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    const int3 n = d->local_config.int3_params[AC_nlocal];
    unsigned long long* slot = d->umax2 + 2 * d->cfl_cur;
    launch_zero_bytes(slot, 2 * sizeof(unsigned long long), s);
    if (d->cfl_mode == 0) {
        const AcReal* uu[3] = {d->in[VTXBUF_UUX], d->in[VTXBUF_UUY], d->in[VTXBUF_UUZ]};
        launch_umax2<AcReal>(uu, n.x, n.y, n.z, slot, s);
    }
    else {
        const AcReal* f[NF];
        for (int q = 0; q < NF; ++q)
            f[q] = d->in[q];
        const double cs = (double)d->local_config.real_params[AC_cs_sound];
        launch_cfl_state<AcReal>(f, d->C, n.x, n.y, n.z, d->cfl_mode, cs * cs, slot, s);
    }
    AC_HIP_CHECK(hipGetLastError());
    d->cfl_courant = (double)courant;
    return AC_SUCCESS;
}

AcResult acDeviceGetTimestep(const Device d, const Stream stream, double* dt)
{
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    AC_HIP_CHECK(hipMemcpyAsync(d->host_pinned, d->umax2 + 2 * d->cfl_cur, 2 * sizeof(double),
                                hipMemcpyDeviceToHost, s));
    AC_HIP_CHECK(hipStreamSynchronize(s));
    SolveArgs<AcReal> a;
    memset(&a, 0, sizeof(a));
    const AcMeshInfo& info = d->local_config;
    a.cfl_cdt = d->cfl_courant;
    a.cfl_dsmin = (double)info.real_params[AC_dsmin];
    a.cfl_cs = (double)info.real_params[AC_cs_sound];
    set_cfl_args(d, a);
    *dt = cfl_dt(a.cfl_mode, d->host_pinned[0], d->host_pinned[1], a.cfl_cdt, a.cfl_dsmin,
                 a.cfl_cs, a.cfl_cs2, a.cfl_cdtv, a.cfl_visc);
    return AC_SUCCESS;
}

AcResult acDeviceIntegrateSubstepDeviceDt(const Device d, const Stream stream,
                                          const int step_number, const AcReal courant)
{
    if (step_number < 0 || step_number > 2)
        return fail("acDeviceIntegrateSubstepDeviceDt: step_number must be 0, 1 or 2");
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    if (ensure_shape(d, s, step_number == 0) != AC_SUCCESS)
        return AC_FAILURE;
    SolveArgs<AcReal> a;
    fill_solve_args(d, a, step_number, (AcReal)0);
    if (d->cfl_mode == 0)
        a.cfl_cdt = (double)courant;
    a.umax2_cur = (const double*)(d->umax2 + 2 * d->cfl_cur);
    const int nxt = 1 - d->cfl_cur;
    if (step_number == 2) {
        unsigned long long* next = d->umax2 + 2 * nxt;
        launch_zero_bytes(next, 2 * sizeof(unsigned long long), s);
        a.time_acc = d->time_acc;
        if (d->cfl_mode == 0) {
            // max |u_new|^2, reduced by the solve kernel itself (by the forcing kernel when
            // forcing follows the solve)
            a.umax2_next = forcing_on(d) ? nullptr : next;
        }
        else {
            // The reference modes reduce |u| and vA of the state entering this substep: one
            // pass over it right before the solve, on the same stream (fusing it into the solve
            // would cost the solve kernels registers they need for two waves per SIMD).
            const AcReal* f[NF];
            for (int q = 0; q < NF; ++q)
                f[q] = d->in[q];
            const int3 n = d->local_config.int3_params[AC_nlocal];
            launch_cfl_state<AcReal>(f, d->C, n.x, n.y, n.z, d->cfl_mode, a.cfl_cs2, next, s);
        }
    }
    const AcResult r = launch_solve(d, a, s);
    if (r == AC_SUCCESS && step_number == 2 && forcing_on(d))
        launch_forcing_step(d, a, s, d->cfl_mode == 0 ? d->umax2 + 2 * nxt : nullptr);
    if (step_number == 2)
        d->cfl_cur = nxt;
    d->cfl_courant = (double)courant;
    return r;
}

AcResult acDeviceIntegrateStepsCFL(const Device d, const Stream stream, const int nsteps,
                                   const AcReal courant, double* time_out)
{
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    const int3 n = d->local_config.int3_params[AC_nlocal];
    // time_acc accumulates the simulated time of this call (the forcing's step start times are
    // AC_current_time plus it)
    launch_zero_bytes(d->time_acc, sizeof(double), s);
    // The tiled solve kernels read the periodic images of the halo cells themselves, so no fill
    // precedes their substeps (launch_solve fills first where another kernel runs). The reference
    // timestep modes reduce over the state entering substep 2 with derivatives that read the
    // halos: that substep keeps its fill. One fill up front also gives the launch autotuner
    // (first substep on a new grid) a state with consistent halos.
    launch_periodic_fill<AcReal>(d->in, NF, n.x, n.y, n.z, s);
    if (ensure_shape(d, s, nsteps > 0) != AC_SUCCESS)
        return AC_FAILURE;
    for (int step = 0; step < nsteps; ++step) {
        for (int sub = 0; sub < 3; ++sub) {
            const bool fused = d->cfl_mode == 0 || sub != 2;
            if (!fused)
                launch_periodic_fill<AcReal>(d->in, NF, n.x, n.y, n.z, s);
            d->periodic_fused = fused;
            d->forcing_time_acc = true;
            const AcResult r = acDeviceIntegrateSubstepDeviceDt(d, stream, sub, courant);
            d->periodic_fused = false;
            d->forcing_time_acc = false;
            if (r != AC_SUCCESS)
                return AC_FAILURE;
            acDeviceSwapBuffers(d);
        }
    }
    AC_HIP_CHECK(hipGetLastError());
    if (time_out) {
        AC_HIP_CHECK(hipMemcpyAsync(d->host_pinned, d->time_acc, sizeof(double), hipMemcpyDeviceToHost, s));
        AC_HIP_CHECK(hipStreamSynchronize(s));
        *time_out += d->host_pinned[0];
    }
    return AC_SUCCESS;
}

AcResult acDeviceReduceDiagnostics(const Device d, const Stream stream, AcDiagnostics* out)
{
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    const int3 n = d->local_config.int3_params[AC_nlocal];
    const AcReal* f[NF];
    for (int q = 0; q < NF; ++q)
        f[q] = d->in[q];
    launch_diagnostics<AcReal>(f, d->C, n.x, n.y, n.z, d->partials, d->max_partials, d->dres + 1, s);
    AC_HIP_CHECK(hipGetLastError());
    AC_HIP_CHECK(hipMemcpyAsync(d->host_pinned, d->dres + 1, sizeof(double) * kDiagSlots,
                                hipMemcpyDeviceToHost, s));
    AC_HIP_CHECK(hipStreamSynchronize(s));
    const double* v = d->host_pinned;
    const double inv_n = 1.0 / ((double)n.x * n.y * n.z);
    for (int q = 0; q < NF; ++q) {
        out->min[q] = v[q];
        out->max[q] = v[NF + q];
        out->rms[q] = sqrt(v[2 * NF + q] * inv_n);
    }
    out->uu_rms = sqrt(v[3 * NF + 0] * inv_n);
    out->uu_max = sqrt(v[3 * NF + 1]);
    out->uu_min = sqrt(v[3 * NF + 2]);
    out->bb_rms = sqrt(v[3 * NF + 3] * inv_n);
    out->bb_max = sqrt(v[3 * NF + 4]);
    out->bb_min = sqrt(v[3 * NF + 5]);
    out->va_rms = sqrt(v[3 * NF + 6] * inv_n);
    out->va_max = sqrt(v[3 * NF + 7]);
    out->va_min = sqrt(v[3 * NF + 8]);
    return AC_SUCCESS;
}

AcResult acDeviceGetVertexBufferPtrs(const Device d, const VertexBufferHandle h, AcReal** in,
                                     AcReal** out, AcReal** w)
{
    if ((int)h < 0 || (int)h >= NF)
        return fail("acDeviceGetVertexBufferPtrs: invalid handle");
    if (in)
        *in = d->in[h];
    if (out)
        *out = d->out[h];
    if (w)
        *w = d->w[h];
    return AC_SUCCESS;
}

AcResult acDeviceSetSolveTiming(const Device d, const int enable)
{
    AC_HIP_CHECK(hipSetDevice(d->id));
    if (enable && !d->events_created) {
        for (int i = 0; i < kTimingSlots; ++i) {
            AC_HIP_CHECK(hipEventCreate(&d->ev_start[i]));
            AC_HIP_CHECK(hipEventCreate(&d->ev_stop[i]));
        }
        d->events_created = true;
    }
    d->timing = enable != 0;
    d->timing_used = 0;
    return AC_SUCCESS;
}

AcResult acDeviceGetSolveTiming(const Device d, double* total_ms, int* launches)
{
    AC_HIP_CHECK(hipSetDevice(d->id));
    AC_HIP_CHECK(hipDeviceSynchronize());
    double tot = 0.0;
    for (int i = 0; i < d->timing_used; ++i) {
        float ms = 0.f;
        AC_HIP_CHECK(hipEventElapsedTime(&ms, d->ev_start[i], d->ev_stop[i]));
        tot += ms;
    }
    if (total_ms)
        *total_ms = tot;
    if (launches)
        *launches = d->timing_used;
    return AC_SUCCESS;
}

AcResult acDeviceGetStreamHandle(const Device d, const Stream stream, void** hip_stream)
{
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    if (!s)
        return fail("acDeviceGetStreamHandle: stream creation failed");
    *hip_stream = (void*)s;
    return AC_SUCCESS;
}

AcResult acDeviceGetSolveLaunchShape(const Device d, int* tile_x, int* tile_y, int* zchunk)
{
    *tile_x = d->shape.tx;
    *tile_y = d->shape.ty;
    *zchunk = d->shape.zchunk;
    return AC_SUCCESS;
}

AcResult acDeviceGetSolveKernel(const Device d, int* kind)
{
    *kind = d->shape.tx <= 0 ? -1 : d->shape.kind;
    return AC_SUCCESS;
}

}  // extern "C"

// ------------------------------------------------------------------ autotuner
// Times every candidate launch (the direct kernel, the delay-line shapes and the sweep shapes) on
// the current grid with the device's own compensation width, physics switches and periodic-image
// reads, and keeps the fastest. No allocation: the trial launches read the real state and write
// f' into the out slab, which holds nothing between substeps. When the tuning runs ahead of a
// substep-0 launch (every caller's first substep), w is dead as well -- that launch overwrites it
// without reading it -- and a trial is one launch of each RK3 substep with w read from and
// written into w, the same bytes and the same substep mix as the stepping (a kernel whose
// substep 0 is fast but whose w-reading substeps are slow was picked when only substep 0 was
// timed). Otherwise (an explicit acDeviceAutotune) a trial is substep 0 alone, which reads f
// and writes f' and w, with w into the out slab too (aliasing f', which halves the trial's write
// traffic). The compensation records (read, and written into the out slab) stay untouched. The choice can be
// cached in the file named by ASTAROTH_AUTOTUNE_CACHE (lines "nx ny nz kahan upwd entropy
// realsize tx ty zchunk kind"); ASTAROTH_AUTOTUNE_VERBOSE=1 prints each candidate's time to
// stderr.
static bool read_cache(Device d, TiledShape* out)
{
    const char* path = getenv("ASTAROTH_AUTOTUNE_CACHE");
    if (!path)
        return false;
    FILE* fp = fopen(path, "r");
    if (!fp)
        return false;
    const int3 n = d->local_config.int3_params[AC_nlocal];
    char line[256];
    bool found = false;
    while (fgets(line, sizeof(line), fp)) {
        int v[11];
        if (sscanf(line, "%d %d %d %d %d %d %d %d %d %d %d", &v[0], &v[1], &v[2], &v[3], &v[4],
                   &v[5], &v[6], &v[7], &v[8], &v[9], &v[10]) != 11)
            continue;
        if (v[0] == n.x && v[1] == n.y && v[2] == n.z && v[3] == d->kahan_bits &&
            v[4] == d->local_config.bool_params[AC_lupwd] &&
            v[5] == d->local_config.bool_params[AC_lentropy] && v[6] == (int)sizeof(AcReal)) {
            const TiledShape s{v[7], v[8], v[9], v[10]};
            if (s.tx == 0 || shape_fits(s, n)) {
                *out = s;
                found = true;
            }
        }
    }
    fclose(fp);
    return found;
}

static void write_cache(Device d, TiledShape s)
{
    const char* path = getenv("ASTAROTH_AUTOTUNE_CACHE");
    if (!path)
        return;
    FILE* fp = fopen(path, "a");
    if (!fp)
        return;
    const int3 n = d->local_config.int3_params[AC_nlocal];
    fprintf(fp, "%d %d %d %d %d %d %d %d %d %d %d\n", n.x, n.y, n.z, d->kahan_bits,
            d->local_config.bool_params[AC_lupwd], d->local_config.bool_params[AC_lentropy],
            (int)sizeof(AcReal), s.tx, s.ty, s.zchunk, s.kind);
    fclose(fp);
}

static const TiledShape kSweepShapes[] = {
    {64, 4, 32, 1}, {64, 4, 64, 1}, {64, 4, 128, 1}, {64, 4, 256, 1},
};

static AcResult autotune(Device d, hipStream_t s, bool w_dead)
{
    const int3 n = d->local_config.int3_params[AC_nlocal];
    TiledShape best{0, 0, 0, 0};
    if (read_cache(d, &best)) {
        d->shape = best;
        d->shape_tuned = true;
        return AC_SUCCESS;
    }
    const bool upwd = d->local_config.bool_params[AC_lupwd] != 0;
    const bool entropy = d->local_config.bool_params[AC_lentropy] != 0;
    const int kb = d->kahan_bits;
    const bool verbose = getenv("ASTAROTH_AUTOTUNE_VERBOSE") != nullptr;
    // Trial outputs: f' into out, w into w when it is dead (else into out), the compensation
    // records (at most NF * KB / 8 <= NF * sizeof(float) bytes per cell) into out. Only a single
    // out slab (no field swapped alone) qualifies; otherwise only the direct kernel is timed.
    // With w dead a trial is one launch of each RK3 substep (s = 1 and 2 read w, s = 0 and 1
    // write it), the mix the stepping runs; with w live, substep 0 alone (it never reads w).
    const int ntrial = w_dead ? 3 : 1;
    SolveArgs<AcReal> ta[3];
    bool one_slab = true;
    for (int q = 0; q < NF; ++q)
        one_slab = one_slab && d->out[q] == d->out[0] + (size_t)q * d->count;
    for (int t = 0; t < ntrial; ++t) {
        fill_solve_args(d, ta[t], t, (AcReal)1e-9);
        for (int q = 0; q < NF; ++q) {
            if (!w_dead)
                ta[t].w[q] = ta[t].out[q];
            ta[t].comp[q] = kb > 0 ? (void*)((char*)d->out[0] + (size_t)q * (kb / 8)) : nullptr;
        }
        ta[t].wrap = 1;
    }
    const SolveArgs<AcReal>& a = ta[0];
    hipEvent_t e0, e1;
    AC_HIP_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return fail("autotune: event creation failed");
    }
    // Each candidate: one warm-up launch, then the best of three timed pairs of launches (the
    // minimum is robust against another process's work or a clock ramp during one sample).
    auto time_candidate = [&](TiledShape sh) -> float {
        auto launch = [&]() -> bool {
            for (int t = 0; t < ntrial; ++t) {
                if (sh.tx <= 0) {
                    SolveArgs<AcReal> ad = ta[t];
                    ad.wrap = 0;  // the direct kernel reads the halos as stored
                    launch_solve_direct<AcReal>(ad, upwd, entropy, kb, s);
                }
                else if (!launch_shape(ta[t], upwd, entropy, kb, sh, s)) {
                    return false;
                }
            }
            return true;
        };
        if (!launch())
            return 1e30f;
        float best_ms = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            if (hipEventRecord(e0, s) != hipSuccess)
                return 1e30f;
            launch();
            launch();
            if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
                return 1e30f;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                return 1e30f;
            best_ms = std::min(best_ms, ms);
        }
        if (verbose)
            fprintf(stderr, "astaroth_amd autotune: %d^3-ish %dx%dx%d kb=%d kind=%d: %.4f ms per launch\n",
                    n.x, sh.tx, sh.ty, sh.zchunk, kb, sh.tx <= 0 ? -1 : sh.kind,
                    best_ms / (2 * ntrial));
        return best_ms;
    };
    std::vector<TiledShape> cands;
    cands.push_back(TiledShape{0, 0, 0, 0});
    if (one_slab) {
        for (int i = 0; i < tiled_num_shapes(); ++i) {
            const TiledShape sh = tiled_shape(i);
            // the ring kernels read halos as stored and never win; they stay pinnable only
            if (sh.tx == 64 && shape_fits(sh, n))
                cands.push_back(sh);
        }
        for (const TiledShape& sh : kSweepShapes)
            if (shape_fits(sh, n) && sweep_applies<AcReal>(a, kb, sh))
                cands.push_back(sh);
    }
    float best_ms = 1e30f;
    for (const TiledShape& sh : cands) {
        const float ms = time_candidate(sh);
        if (ms < best_ms) {
            best_ms = ms;
            best = sh;
        }
    }
    const hipError_t err = hipGetLastError();
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (err != hipSuccess)
        return fail("autotune: a trial launch failed");
    d->shape = best;
    d->shape_tuned = true;
    write_cache(d, best);
    return AC_SUCCESS;
}

// w_dead: the caller's next launch is substep 0, which writes w without reading it.
static AcResult ensure_shape(Device d, hipStream_t s, bool w_dead)
{
    if (d->shape_tuned)
        return AC_SUCCESS;
    return autotune(d, s, w_dead);
}

extern "C" AcResult acDeviceAutotune(const Device d, const Stream stream)
{
    AC_HIP_CHECK(hipSetDevice(d->id));
    d->shape_tuned = false;
    return autotune(d, get_stream(d, stream), false);
}

// ------------------------------------------------------------------ helpers for adjoint.hip
namespace acamd {
void device_solve_args(Device d, SolveArgs<AcReal>& a, int step, AcReal dt)
{
    fill_solve_args(d, a, step, dt);
}
hipStream_t device_stream(Device d, Stream s) { return get_stream(d, s); }
int device_id(Device d) { return d->id; }
}  // namespace acamd

extern "C" AcResult acDeviceLoadMeshFromDevicePtr(const Device d, const Stream stream,
                                                  const AcReal* src)
{
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    for (int q = 0; q < NF; ++q)
        AC_HIP_CHECK(hipMemcpyAsync(d->in[q], src + (size_t)q * d->count, d->count * sizeof(AcReal),
                                    hipMemcpyDeviceToDevice, s));
    if (d->kahan_bits > 0)
        launch_zero_bytes(d->slab_comp, d->count * (size_t)(d->kahan_bits / 8) * NF, s);
    return AC_SUCCESS;
}

extern "C" AcResult acDeviceStoreMeshToDevicePtr(const Device d, const Stream stream, AcReal* dst)
{
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    for (int q = 0; q < NF; ++q)
        AC_HIP_CHECK(hipMemcpyAsync(dst + (size_t)q * d->count, d->in[q], d->count * sizeof(AcReal),
                                    hipMemcpyDeviceToDevice, s));
    return AC_SUCCESS;
}

// ------------------------------------------------------------------ snapshots on disk
// The reference's snapshot format (ref src/core/grid.cc:4405-4530 with collective IO, and
// samples/standalone_mpi/main.cc:155-216): one raw file per evolved field,
// "<dir>/<field name>-<label>.mesh", holding the interior cells only (nx * ny * nz AcReal values,
// x fastest), plus the restart table snapshots_info.csv written by acHostAppendSnapshotInfo.
extern "C" const char* acGetVertexBufferName(const VertexBufferHandle handle);

static std::string snapshot_path(const char* dir, int q, const char* label)
{
    return std::string(dir) + "/" + acGetVertexBufferName((VertexBufferHandle)q) + "-" + label +
           ".mesh";
}

extern "C" AcResult acDeviceWriteSnapshot(const Device d, const Stream stream, const char* dir,
                                          const char* label)
{
    if (!dir || !label)
        return fail("acDeviceWriteSnapshot: dir and label are required");
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    const int3 n = d->local_config.int3_params[AC_nlocal];
    const size_t mx = (size_t)n.x + 2 * R, my = (size_t)n.y + 2 * R;
    const size_t plane = (size_t)n.x * n.y;
    AcReal* host = nullptr;
    AC_HIP_CHECK(hipHostMalloc((void**)&host, plane * n.z * sizeof(AcReal)));
    AcResult r = AC_SUCCESS;
    for (int q = 0; q < NF && r == AC_SUCCESS; ++q) {
        for (int k = 0; k < n.z; ++k) {
            const AcReal* src = d->in[q] + R + mx * (R + my * (size_t)(k + R));
            if (hipMemcpy2DAsync(host + plane * k, n.x * sizeof(AcReal), src, mx * sizeof(AcReal),
                                 n.x * sizeof(AcReal), n.y, hipMemcpyDeviceToHost, s) != hipSuccess) {
                r = fail("acDeviceWriteSnapshot: device-to-host copy failed");
                break;
            }
        }
        if (r != AC_SUCCESS || hipStreamSynchronize(s) != hipSuccess) {
            r = AC_FAILURE;
            break;
        }
        const std::string path = snapshot_path(dir, q, label);
        FILE* fp = fopen(path.c_str(), "wb");
        if (!fp) {
            r = fail("acDeviceWriteSnapshot: cannot open a snapshot file for writing");
            break;
        }
        const size_t cnt = plane * n.z;
        if (fwrite(host, sizeof(AcReal), cnt, fp) != cnt)
            r = fail("acDeviceWriteSnapshot: short write");
        fclose(fp);
    }
    hipHostFree(host);
    return r;
}

extern "C" AcResult acDeviceReadSnapshot(const Device d, const Stream stream, const char* dir,
                                         const char* label)
{
    if (!dir || !label)
        return fail("acDeviceReadSnapshot: dir and label are required");
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = get_stream(d, stream);
    const int3 n = d->local_config.int3_params[AC_nlocal];
    const size_t mx = (size_t)n.x + 2 * R, my = (size_t)n.y + 2 * R;
    const size_t plane = (size_t)n.x * n.y;
    const size_t cnt = plane * n.z;
    AcReal* host = nullptr;
    AC_HIP_CHECK(hipHostMalloc((void**)&host, cnt * sizeof(AcReal)));
    AcResult r = AC_SUCCESS;
    for (int q = 0; q < NF && r == AC_SUCCESS; ++q) {
        const std::string path = snapshot_path(dir, q, label);
        FILE* fp = fopen(path.c_str(), "rb");
        if (!fp) {
            r = fail("acDeviceReadSnapshot: cannot open a snapshot file");
            break;
        }
        const size_t got = fread(host, sizeof(AcReal), cnt, fp);
        const bool extra = fgetc(fp) != EOF;
        fclose(fp);
        if (got != cnt || extra) {
            r = fail("acDeviceReadSnapshot: snapshot file size does not match the grid");
            break;
        }
        for (int k = 0; k < n.z; ++k) {
            AcReal* dst = d->in[q] + R + mx * (R + my * (size_t)(k + R));
            if (hipMemcpy2DAsync(dst, mx * sizeof(AcReal), host + plane * k, n.x * sizeof(AcReal),
                                 n.x * sizeof(AcReal), n.y, hipMemcpyHostToDevice, s) != hipSuccess) {
                r = fail("acDeviceReadSnapshot: host-to-device copy failed");
                break;
            }
        }
        // the pinned buffer is reused for the next field: wait for this one's copies
        if (hipStreamSynchronize(s) != hipSuccess)
            r = AC_FAILURE;
    }
    hipHostFree(host);
    if (r != AC_SUCCESS)
        return r;
    launch_periodic_fill<AcReal>(d->in, NF, n.x, n.y, n.z, s);
    if (d->kahan_bits > 0)
        launch_zero_bytes(d->slab_comp, d->count * (size_t)(d->kahan_bits / 8) * NF, s);
    AC_HIP_CHECK(hipGetLastError());
    return AC_SUCCESS;
}
