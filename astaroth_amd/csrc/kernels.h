// kernels.h -- launchers of the HIP kernels of the MHD hot path, called by the device layer.
#pragma once

#include <hip/hip_runtime.h>

#include "mhd_common.h"

#ifndef AC_DOUBLE_PRECISION
#define AC_DOUBLE_PRECISION 0
#endif

namespace acamd {

#if AC_DOUBLE_PRECISION
typedef double AcRealT;
#else
typedef float AcRealT;
#endif

// Wave-level max of a non-negative double, one atomic per wave64 into `dst`.
// Every lane of the calling wave must reach this call.
__device__ __forceinline__ void wave_max_to_global(double v, unsigned long long* dst)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = fmax(v, __shfl_xor(v, off, 64));
    const unsigned lane = (threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z)) & 63u;
    if (lane == 0u)
        atomic_max_nonneg(dst, v);
}

// ---- solve ----
template <typename T>
void launch_solve_direct(const SolveArgs<T>& a, bool upwd, bool entropy, int kahan_bits,
                         hipStream_t stream);

// LDS-tiled 2.5D solve. Returns false when the launch shape does not fit the grid.
// Tiles of 64 x 4 and 64 x 2 run the delay-line kernel (solve_tiled.hip); tiles of 32 x 8,
// 16 x 16 and 32 x 16 the four-plane ring kernel (solve_ring.hip).
struct TiledShape {
    int tx, ty, zchunk;
    int kind;  // 0: delay-line / ring kernel (solve_tiled.hip, solve_ring.hip); 1: sweep (solve_sweep.hip)
};
template <typename T>
bool launch_solve_tiled(const SolveArgs<T>& a, bool upwd, bool entropy, int kahan_bits,
                        TiledShape shape, hipStream_t stream);
template <typename T>
bool launch_solve_ring(const SolveArgs<T>& a, bool upwd, bool entropy, int kahan_bits,
                       TiledShape shape, hipStream_t stream);
// The unrolled sweep kernel (FP32, 64 x 4 column blocks, grids with equal spacing along the three
// axes). Returns false (nothing launched) when it does not apply.
template <typename T>
bool launch_solve_sweep(const SolveArgs<T>& a, bool upwd, bool entropy, int kahan_bits,
                        TiledShape shape, hipStream_t stream);
bool sweep_shape_fits(TiledShape s, int nx, int ny, int nz);
// Whether launch_solve_sweep runs the sweep kernel for these arguments (it honours a.wrap).
template <typename T>
bool sweep_applies(const SolveArgs<T>& a, int kahan_bits, TiledShape shape);
// The adjoint's local stage on the tiled machinery: derivative adjoints g (NG fields of m^3,
// same field stride as the state), lambda_f / lambda_w (NF fields, updated in place at the
// interior cells: lambda_w <- alpha mu, lambda_f <- lambda_f + dt * the value adjoints), (nu, eta)
// block partials.
template <typename T>
struct AdjLocal {
    T* g;
    T* lf;
    T* lw;
    double* partials;
};
// Returns false (nothing launched) when the grid does not fit 64 x 4 tiles.
template <typename T>
bool launch_adjoint_local_tiled(const SolveArgs<T>& a, const AdjLocal<T>& x, bool upwd,
                                bool entropy, hipStream_t s, long long* nblocks);

// LDS pointer type of the global->LDS DMA builtin
typedef __attribute__((address_space(3))) void* lds_ptr_t;

// End of a plane step of the LDS-staged z-marching kernels: wait for this wave's memory
// operations (its DMA into the other plane buffer among them), then the workgroup barrier.
// The global->LDS DMA builtin writes LDS at M0 + inst_offset + 4 * lane, a range the optimiser
// does not see (its pointer operand is the wave's base). Without the memory clobbers around the
// barrier it may keep an LDS value read before the barrier in a register for a later read of
// the same address, or issue a read after the barrier, when another wave's DMA of the next
// plane can already have overwritten it; the clobbers pin every LDS access to its own side.
__device__ __forceinline__ void plane_barrier()
{
    __builtin_amdgcn_s_waitcnt(0);
    asm volatile("" ::: "memory");
    __syncthreads();
    asm volatile("" ::: "memory");
}

// Returns v unchanged, but opaque to the optimiser. Used on per-lane 32-bit offsets inside the
// plane loops: otherwise loop-invariant code motion folds them into 64-bit per-lane addresses
// of every field array (a VGPR pair per field and array) that stay live across the loop, where
// a uniform 64-bit base in SGPRs plus the 32-bit offset is what the memory instructions take.
__device__ __forceinline__ unsigned opaque(unsigned v)
{
    asm volatile("" : "+v"(v));
    return v;
}
// The same for a wave-uniform 64-bit value kept in a scalar register pair: recomputing
// per-field bases from it each plane step beats holding eight bases per array across the loop.
__device__ __forceinline__ long long opaque_s64(long long v)
{
    asm volatile("" : "+s"(v));
    return v;
}
// A per-field array base made to live in a scalar register pair: a global access then takes
// the form uniform base + 32-bit lane offset instead of a per-lane 64-bit address built with
// vector adds, which the compiler otherwise derives field by field from the previous one.
__device__ __forceinline__ gchar* sbase(const void* p)
{
    return (gchar*)opaque_s64((long long)p);
}
// z-chunk of the adjoint's local stage (64 x 4 column blocks, one 256-thread workgroup per CU):
// long chunks amortise the six plane steps of each chunk's start-up; the longest of 256, 128
// and 64 (up to all of nz) that still gives at least 256 workgroups, one per CU of the MI355X
// (256^3: one 256-plane chunk per column, 9.20 -> 9.01 ms per adjoint step against 128-plane
// chunks), else the largest of 32, 16 and 8 below nz.
inline int adjoint_zchunk(int nx, int ny, int nz)
{
    const long long tiles = (long long)(nx / 64) * (ny / 4);
    for (int cand : {256, 128, 64})
        if (nz % cand == 0 && nz >= cand && tiles * (nz / cand) >= 256)
            return cand;
    for (int cand : {32, 16, 8})
        if (nz % cand == 0 && nz > cand)
            return cand;
    return nz;
}
// Candidate launch shapes of the tiled kernels (for the autotuner).
int tiled_num_shapes();
TiledShape tiled_shape(int i);
bool tiled_shape_fits(TiledShape s, int nx, int ny, int nz);
// Whether launch_solve_tiled with this shape runs a delay-line kernel that honours a.wrap (reads
// the periodic images instead of the halos).
template <typename T>
bool tiled_reads_periodic(const SolveArgs<T>& a, int kb, TiledShape sh);

// ---- seeded Fourier-mode initial condition (ic.hip) ----
template <typename T>
struct FourierArgs {
    T* uu[3];
    T* aa[3];
    T* lnrho;
    T* ss;
    const double* head;  // per (group, mode): cos phase, sin phase, ampl * e (3)
    const double* tab;   // per (group, mode): cos / sin of k_a x_a along x, y, z
    int nmodes;
    int nx, ny, nz, mx, my, mz;
    double lnrho_perturb;
    unsigned long long lseed;
};
template <typename T>
void launch_fourier_ic(const FourierArgs<T>& A, hipStream_t s);

// ---- inputs of the reference timestep (cfl_cell modes 1, 2) for the current state ----
template <typename T>
void launch_cfl_state(const T* const* fields, const Coeffs<T>& C, int nx, int ny, int nz, int mode,
                      double cs2, unsigned long long* out, hipStream_t s);

// ---- one-pass diagnostics: per field min, max, sum of squares; then sum, max and min of
// |u|^2, |curl A|^2 and |curl A|^2 / exp(lnrho)
// Zeroes field q's entries of the per-cell compensation records (count cells, NF entries of
// kahan_bits / 8 bytes each).
void launch_zero_comp_field(void* slab, size_t count, int kahan_bits, int q, hipStream_t stream);

// ---- helical forcing (aux_kernels.hip) ----
// Adds the forcing of ref mhdsolver.ac:286-331 to u (the `out` buffers of the substep-2 solve,
// interior cells) when the step's start time exceeds switch_forcing; with umax2_next set, also
// reduces max |u|^2 of the forced state there (the courant timestep mode, in place of the
// solve's own reduction).
template <typename T>
struct ForcingArgs {
    T* uu[3];
    int nx, ny, nz, mx, my;
    long long sz;
    int gx0, gy0, gz0;               // global index of the first interior cell (0 on one device)
    int ngx, ngy, ngz;               // AC_ngrid
    double ds[3];
    double k[3], fre[3], fim[3], phase;
    double nn;                       // cs AC_forcing_magnitude sqrt(AC_kaver cs)
    double t_base, switch_time;      // start time of the step = t_base [+ *time_acc - dt]
    const double* time_acc;          // device-accumulated time including this step, or null
    SolveArgs<T> s;                  // dt of the step (solve_dt)
    unsigned long long* umax2_next;  // or null
};
template <typename T>
void launch_forcing(const ForcingArgs<T>& a, hipStream_t stream);

constexpr int DIAG_SLOTS = 3 * NF + 9;

// ---- periodic fill: every halo cell of every listed field, one launch ----
template <typename T>
void launch_periodic_fill(T* const* fields, int nfields, int nx, int ny, int nz, hipStream_t s);
// The halo cells of the box [lo, hi) of the m^3 arrays (sub-volume fill); interior cells untouched.
template <typename T>
void launch_periodic_fill_box(T* const* fields, int nfields, int nx, int ny, int nz, const int* lo,
                              const int* hi, hipStream_t s);

// ---- reductions ----
// map: AcMapKernel value; op: AcReduceOp value. Interior cells only. Result is the raw
// FP64 reduction (no RMS post-processing), written to *result_dev (device memory).
template <typename T>
void launch_reduce(const T* const* bufs, int map, int op, int nx, int ny, int nz,
                   double* partials, int max_partials, double* result_dev, hipStream_t s);
int reduce_num_partials(int nx, int ny, int nz);

// One-pass diagnostics: writes 4 * NF + 4 doubles to out_dev:
// [min x8][max x8][sum sq x8][sum |u|^2][max |u|^2][sum |B|^2][max |B|^2]
template <typename T>
void launch_diagnostics(const T* const* fields, const Coeffs<T>& C, int nx, int ny, int nz,
                        double* partials, int max_partials, double* out_dev, hipStream_t s);

// max |u|^2 over the interior into umax2 (atomic max, caller zeroes it first).
template <typename T>
void launch_umax2(const T* const* uu, int nx, int ny, int nz, unsigned long long* umax2,
                  hipStream_t s);

// Fill n values at dst with v.
template <typename T>
void launch_fill(T* dst, size_t n, T v, hipStream_t s);
void launch_zero_bytes(void* dst, size_t nbytes, hipStream_t s);

}  // namespace acamd
