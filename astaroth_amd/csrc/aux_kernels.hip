// aux_kernels.hip -- periodic halo fill (one launch for all fields), FP64-accumulating
// reductions, one-pass diagnostics and small utility kernels.
//
// Periodic fill: ref acc-runtime/built-in/kernels.h:17-41 (BOUNDCOND_PERIODIC_DEVICE) launched
// per field over the whole m^3 volume by ref src/core/device.cc:1009-1017. Here one launch
// enumerates only the m^3 - n^3 halo cells of every field; the source is always an interior
// cell, so the fill has no read/write hazards.
// Reductions: ref src/core/kernels/reductions.cu:98-219 (map kernel into an n^3 scratchpad, then
// hipcub DeviceReduce). Here the map is fused into a single pass with FP64 accumulation and a
// deterministic two-stage tree (fixed grid, fixed order).
#include <cstdio>
#include <cstdlib>

#include "kernels.h"

namespace acamd {

constexpr int RED_TPB = 256;
constexpr int RED_MAX_BLOCKS = 1024;

// ---------------------------------------------------------------- periodic fill
template <typename T>
struct FillArgs {
    T* f[NF];
    int nx, ny, nz, mx, my;
    unsigned halo_count;  // per field (< 2^32: checked at launch)
    int wide;             // every extent >= R: a halo cell's image is one period away
};

// One thread per halo cell and field. The cell enumeration (two z slabs of 3 full planes, then
// per interior plane 3 + 3 full rows and 3 + 3 cells at the ends of each interior row) is
// inverted with 32-bit divisions, and the image is the cell one period away (wrapped with a
// true modulo only on grids narrower than the halo), so the kernel is bound by its memory
// traffic rather than by 64-bit integer division.
template <typename T>
__global__ void __launch_bounds__(256) periodic_fill_kernel(FillArgs<T> a)
{
    const unsigned t0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (t0 >= a.halo_count)
        return;
    T* __restrict__ f = a.f[blockIdx.y];
    const unsigned mx = (unsigned)a.mx, plane = mx * (unsigned)a.my;
    int i, j, k;
    unsigned t = t0;
    if (t < 6u * plane) {
        const unsigned kk = t / plane;
        const unsigned r = t - kk * plane;
        k = kk < 3u ? (int)kk : a.nz + (int)kk;  // planes 0..2, then nz + 3 .. nz + 5
        j = (int)(r / mx);
        i = (int)(r - (unsigned)j * mx);
    }
    else {
        t -= 6u * plane;
        const unsigned per = 6u * mx + 6u * (unsigned)a.ny;
        const unsigned kk = t / per;
        unsigned r = t - kk * per;
        k = R + (int)kk;
        if (r < 6u * mx) {
            const unsigned jj = r / mx;
            j = jj < 3u ? (int)jj : a.ny + (int)jj;  // rows 0..2, then ny + 3 .. ny + 5
            i = (int)(r - jj * mx);
        }
        else {
            r -= 6u * mx;
            const unsigned row = r / 6u;
            const int cc = (int)(r - row * 6u);
            j = R + (int)row;
            i = cc < 3 ? cc : a.nx + cc;
        }
    }
    int is, js, ks;
    if (a.wide) {
        is = i < R ? i + a.nx : (i >= a.nx + R ? i - a.nx : i);
        js = j < R ? j + a.ny : (j >= a.ny + R ? j - a.ny : j);
        ks = k < R ? k + a.nz : (k >= a.nz + R ? k - a.nz : k);
    }
    else {  // proper modulo: an extent smaller than the halo width wraps more than once
        is = ((i - R) % a.nx + a.nx) % a.nx + R;
        js = ((j - R) % a.ny + a.ny) % a.ny + R;
        ks = ((k - R) % a.nz + a.nz) % a.nz + R;
    }
    const long long pl = (long long)plane;
    const long long dst = (long long)i + (long long)a.mx * j + pl * k;
    const long long src = (long long)is + (long long)a.mx * js + pl * ks;
    f[dst] = f[src];
}

template <typename T>
void launch_periodic_fill(T* const* fields, int nfields, int nx, int ny, int nz, hipStream_t s)
{
    FillArgs<T> a;
    for (int q = 0; q < NF; ++q)
        a.f[q] = q < nfields ? fields[q] : nullptr;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.mx = nx + 2 * R;
    a.my = ny + 2 * R;
    const long long hc = 6LL * a.mx * a.my + (long long)nz * (6LL * a.mx + 6LL * ny);
    if (hc >= (1LL << 32) || 6LL * a.mx * a.my >= (1LL << 32)) {
        fprintf(stderr, "astaroth_amd: periodic fill: %lld halo cells per field exceed 2^32\n", hc);
        abort();
    }
    a.halo_count = (unsigned)hc;
    a.wide = nx >= R && ny >= R && nz >= R;
    const dim3 tpb(256, 1, 1);
    const dim3 bpg((unsigned)((a.halo_count + 255) / 256), (unsigned)nfields, 1);
    hipLaunchKernelGGL(periodic_fill_kernel<T>, bpg, tpb, 0, s, a);
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double red_init(int op)
{
    return op == 1 ? INFINITY : (op == 2 ? -INFINITY : 0.0);
}
__device__ __forceinline__ double red_combine(int op, double a, double b)
{
    return op == 1 ? fmin(a, b) : (op == 2 ? fmax(a, b) : a + b);
}

template <typename T>
__device__ __forceinline__ double red_map(int map, const T* const* bufs, long long c)
{
    switch (map) {
    case 1: return (double)bufs[0][c];
    case 2: { const double v = bufs[0][c]; return v * v; }
    case 3: { const double e = exp((double)bufs[0][c]); return e * e; }
    case 4: {
        const double x = bufs[0][c], y = bufs[1][c], z = bufs[2][c];
        return sqrt(x * x + y * y + z * z);
    }
    case 5: {
        const double x = bufs[0][c], y = bufs[1][c], z = bufs[2][c];
        return x * x + y * y + z * z;
    }
    case 6: {
        const double x = exp((double)bufs[0][c]), y = exp((double)bufs[1][c]),
                     z = exp((double)bufs[2][c]);
        return x * x + y * y + z * z;
    }
    case 7: {
        const double x = bufs[0][c], y = bufs[1][c], z = bufs[2][c];
        return sqrt(x * x + y * y + z * z) / sqrt(exp((double)bufs[3][c]));
    }
    case 8: {
        const double x = bufs[0][c], y = bufs[1][c], z = bufs[2][c];
        return (x * x + y * y + z * z) / exp((double)bufs[3][c]);
    }
    default: return 0.0;
    }
}

// Interior cell number t (x fastest) -> index in an m^3 array, in the index type I: 32-bit
// unsigned when the interior has fewer than 2^31 cells and the array fewer than 2^32 (every grid
// up to about 1600^3; use_index32 below), 64-bit otherwise. The grid-stride loops visit the same
// cells in the same order with either type, so the reductions are bit-identical.
template <typename I>
__device__ __forceinline__ I interior_index(I t, I nx, I ny, I mx, I sz)
{
    const I i = t % nx;
    const I r = t / nx;
    const I j = r % ny;
    const I k = r / ny;
    return (i + R) + mx * (j + R) + sz * (k + R);
}

static bool use_index32(int nx, int ny, int nz)
{
    const long long n = (long long)nx * ny * nz;
    const long long m = (long long)(nx + 2 * R) * (ny + 2 * R) * (nz + 2 * R);
    return n < (1ll << 31) && m < (1ll << 32);
}

// The halo cells inside a box [x0, x1) x [y0, y1) x [z0, z1) of the m^3 array (a sub-volume fill,
// ref src/core/device.cc:998-1017 with the BOUNDCOND_PERIODIC_DEVICE kernel of
// acc-runtime/built-in/kernels.h:17-41, which skips the interior): one thread per box cell and
// field; interior cells are left alone, halo cells take their periodic image.
template <typename T>
struct FillBoxArgs {
    T* f[NF];
    int nx, ny, nz, mx, my;
    int x0, y0, z0, bx, by;
    unsigned count;  // box cells per field
};

template <typename T>
__global__ void __launch_bounds__(256) periodic_fill_box_kernel(FillBoxArgs<T> a)
{
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.count)
        return;
    const unsigned bx = (unsigned)a.bx, by = (unsigned)a.by;
    const unsigned r = t / bx;
    const int i = a.x0 + (int)(t - r * bx);
    const int j = a.y0 + (int)(r % by);
    const int k = a.z0 + (int)(r / by);
    const bool inside = i >= R && i < a.nx + R && j >= R && j < a.ny + R && k >= R && k < a.nz + R;
    if (inside)
        return;
    const int is = ((i - R) % a.nx + a.nx) % a.nx + R;
    const int js = ((j - R) % a.ny + a.ny) % a.ny + R;
    const int ks = ((k - R) % a.nz + a.nz) % a.nz + R;
    T* __restrict__ f = a.f[blockIdx.y];
    const long long pl = (long long)a.mx * a.my;
    f[(long long)i + (long long)a.mx * j + pl * k] = f[(long long)is + (long long)a.mx * js + pl * ks];
}

template <typename T>
void launch_periodic_fill_box(T* const* fields, int nfields, int nx, int ny, int nz, const int* lo,
                              const int* hi, hipStream_t s)
{
    FillBoxArgs<T> a;
    for (int q = 0; q < NF; ++q)
        a.f[q] = q < nfields ? fields[q] : nullptr;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.mx = nx + 2 * R;
    a.my = ny + 2 * R;
    a.x0 = lo[0];
    a.y0 = lo[1];
    a.z0 = lo[2];
    a.bx = hi[0] - lo[0];
    a.by = hi[1] - lo[1];
    const long long cnt = (long long)a.bx * a.by * (hi[2] - lo[2]);
    if (cnt <= 0)
        return;
    a.count = (unsigned)cnt;  // a box of an m^3 array the caller checked (< 2^32 cells)
    const dim3 tpb(256, 1, 1);
    const dim3 bpg((unsigned)((cnt + 255) / 256), (unsigned)nfields, 1);
    hipLaunchKernelGGL(periodic_fill_box_kernel<T>, bpg, tpb, 0, s, a);
}

template <typename T>
struct RedArgs {
    const T* bufs[4];
    int map, op;
    int nx, ny, nz, mx, my;
    double* partials;
};

__device__ __forceinline__ double block_reduce(int op, double v, double* sh)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = red_combine(op, v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        sh[wid] = v;
    __syncthreads();
    double r = red_init(op);
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int q = 0; q < nw; ++q)
            r = red_combine(op, r, sh[q]);
    }
    return r;
}

template <typename T, typename I>
__global__ void __launch_bounds__(RED_TPB) reduce_stage1(RedArgs<T> a)
{
    __shared__ double sh[RED_TPB / 64];
    const I n = (I)a.nx * (I)a.ny * (I)a.nz;
    const I mx = (I)a.mx, sz = (I)a.mx * (I)a.my;
    double acc = red_init(a.op);
    for (I t = (I)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (I)gridDim.x * blockDim.x) {
        const I c = interior_index<I>(t, (I)a.nx, (I)a.ny, mx, sz);
        acc = red_combine(a.op, acc, red_map(a.map, a.bufs, (long long)c));
    }
    const double r = block_reduce(a.op, acc, sh);
    if (threadIdx.x == 0)
        a.partials[blockIdx.x] = r;
}

__global__ void __launch_bounds__(RED_TPB) reduce_stage2(const double* partials, int np, int op,
                                                         double* out)
{
    __shared__ double sh[RED_TPB / 64];
    double acc = red_init(op);
    for (int t = threadIdx.x; t < np; t += blockDim.x)
        acc = red_combine(op, acc, partials[t]);
    const double r = block_reduce(op, acc, sh);
    if (threadIdx.x == 0)
        *out = r;
}

int reduce_num_partials(int nx, int ny, int nz)
{
    const long long n = (long long)nx * ny * nz;
    long long b = (n + RED_TPB - 1) / RED_TPB;
    if (b > RED_MAX_BLOCKS)
        b = RED_MAX_BLOCKS;
    if (b < 1)
        b = 1;
    return (int)b;
}

template <typename T>
void launch_reduce(const T* const* bufs, int map, int op, int nx, int ny, int nz, double* partials,
                   int max_partials, double* result_dev, hipStream_t s)
{
    RedArgs<T> a;
    for (int q = 0; q < 4; ++q)
        a.bufs[q] = bufs[q];
    a.map = map;
    a.op = op;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.mx = nx + 2 * R;
    a.my = ny + 2 * R;
    a.partials = partials;
    int nb = reduce_num_partials(nx, ny, nz);
    if (nb > max_partials)
        nb = max_partials;
    if (use_index32(nx, ny, nz))
        hipLaunchKernelGGL((reduce_stage1<T, unsigned>), dim3(nb), dim3(RED_TPB), 0, s, a);
    else
        hipLaunchKernelGGL((reduce_stage1<T, unsigned long long>), dim3(nb), dim3(RED_TPB), 0, s, a);
    hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(RED_TPB), 0, s, (const double*)partials, nb,
                       op, result_dev);
}

// ---------------------------------------------------------------- diagnostics

template <typename T>
struct DiagArgs {
    const T* f[NF];
    Coeffs<T> C;
    int nx, ny, nz, mx, my;
    double* partials;  // DIAG_SLOTS * gridDim.x
};

__device__ __forceinline__ int diag_op(int slot)
{
    // [min x8][max x8][sum sq x8] then, for |u|^2, |B|^2 and vA^2 = |B|^2 / rho: [sum][max][min]
    if (slot < NF)
        return 1;
    if (slot < 2 * NF)
        return 2;
    if (slot < 3 * NF)
        return 3;
    const int r = (slot - 3 * NF) % 3;
    return r == 0 ? 3 : (r == 1 ? 2 : 1);
}

template <typename T, typename I>
__global__ void __launch_bounds__(RED_TPB) diagnostics_stage1(DiagArgs<T> a)
{
    __shared__ double sh[RED_TPB / 64];
    double acc[DIAG_SLOTS];
#pragma unroll
    for (int q = 0; q < DIAG_SLOTS; ++q)
        acc[q] = red_init(diag_op(q));
    const I n = (I)a.nx * (I)a.ny * (I)a.nz;
    const I sy = (I)a.mx, sz = (I)a.mx * (I)a.my;
    for (I t = (I)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (I)gridDim.x * blockDim.x) {
        const I c = interior_index<I>(t, (I)a.nx, (I)a.ny, sy, sz);
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            const double v = a.f[q][c];
            acc[q] = fmin(acc[q], v);
            acc[NF + q] = fmax(acc[NF + q], v);
            acc[2 * NF + q] += v * v;
        }
        const double ux = a.f[1][c], uy = a.f[2][c], uz = a.f[3][c];
        const double u2 = ux * ux + uy * uy + uz * uz;
        acc[3 * NF + 0] += u2;
        acc[3 * NF + 1] = fmax(acc[3 * NF + 1], u2);
        acc[3 * NF + 2] = fmin(acc[3 * NF + 2], u2);
        const I st[3] = {1, sy, sz};
        double g[3][3];
#pragma unroll
        for (int comp = 0; comp < 3; ++comp) {
            const T* f = a.f[4 + comp];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const I s = st[ax];
                const T* kk = a.C.d1[ax];
                g[comp][ax] = (double)(kk[1] * (f[c + s] - f[c - s]) + kk[2] * (f[c + 2 * s] - f[c - 2 * s]) +
                                       kk[3] * (f[c + 3 * s] - f[c - 3 * s]));
            }
        }
        const double bx = g[2][1] - g[1][2], by = g[0][2] - g[2][0], bz = g[1][0] - g[0][1];
        const double b2 = bx * bx + by * by + bz * bz;
        acc[3 * NF + 3] += b2;
        acc[3 * NF + 4] = fmax(acc[3 * NF + 4], b2);
        acc[3 * NF + 5] = fmin(acc[3 * NF + 5], b2);
        // Alfven speed |B| / sqrt(rho), rho = exp(lnrho) (ref acc-runtime/stdlib/map.h:121-134)
        const double va2 = b2 / exp((double)a.f[0][c]);
        acc[3 * NF + 6] += va2;
        acc[3 * NF + 7] = fmax(acc[3 * NF + 7], va2);
        acc[3 * NF + 8] = fmin(acc[3 * NF + 8], va2);
    }
    for (int q = 0; q < DIAG_SLOTS; ++q) {
        const double r = block_reduce(diag_op(q), acc[q], sh);
        if (threadIdx.x == 0)
            a.partials[(long long)q * gridDim.x + blockIdx.x] = r;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(RED_TPB) diagnostics_stage2(const double* partials, int np, double* out)
{
    __shared__ double sh[RED_TPB / 64];
    for (int q = 0; q < DIAG_SLOTS; ++q) {
        const int op = diag_op(q);
        double acc = red_init(op);
        for (int t = threadIdx.x; t < np; t += blockDim.x)
            acc = red_combine(op, acc, partials[(long long)q * np + t]);
        const double r = block_reduce(op, acc, sh);
        if (threadIdx.x == 0)
            out[q] = r;
        __syncthreads();
    }
}

template <typename T>
void launch_diagnostics(const T* const* fields, const Coeffs<T>& C, int nx, int ny, int nz,
                        double* partials, int max_partials, double* out_dev, hipStream_t s)
{
    DiagArgs<T> a;
    for (int q = 0; q < NF; ++q)
        a.f[q] = fields[q];
    a.C = C;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.mx = nx + 2 * R;
    a.my = ny + 2 * R;
    a.partials = partials;
    int nb = reduce_num_partials(nx, ny, nz);
    if (nb * DIAG_SLOTS > max_partials)
        nb = max_partials / DIAG_SLOTS;
    if (use_index32(nx, ny, nz))
        hipLaunchKernelGGL((diagnostics_stage1<T, unsigned>), dim3(nb), dim3(RED_TPB), 0, s, a);
    else
        hipLaunchKernelGGL((diagnostics_stage1<T, unsigned long long>), dim3(nb), dim3(RED_TPB), 0,
                           s, a);
    hipLaunchKernelGGL(diagnostics_stage2, dim3(1), dim3(RED_TPB), 0, s, (const double*)partials,
                       nb, out_dev);
}

// ---------------------------------------------------------------- max |u|^2
template <typename T>
struct UmaxArgs {
    const T* u[3];
    int nx, ny, nz, mx, my;
    unsigned long long* umax2;
};

template <typename T, typename I>
__global__ void __launch_bounds__(RED_TPB) umax2_kernel(UmaxArgs<T> a)
{
    const I n = (I)a.nx * (I)a.ny * (I)a.nz;
    const I mx = (I)a.mx, sz = (I)a.mx * (I)a.my;
    double m = 0.0;
    for (I t = (I)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (I)gridDim.x * blockDim.x) {
        const I c = interior_index<I>(t, (I)a.nx, (I)a.ny, mx, sz);
        const double x = a.u[0][c], y = a.u[1][c], z = a.u[2][c];
        m = fmax(m, x * x + y * y + z * z);
    }
    wave_max_to_global(m, a.umax2);
}

template <typename T>
void launch_umax2(const T* const* uu, int nx, int ny, int nz, unsigned long long* umax2,
                  hipStream_t s)
{
    UmaxArgs<T> a;
    for (int q = 0; q < 3; ++q)
        a.u[q] = uu[q];
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.mx = nx + 2 * R;
    a.my = ny + 2 * R;
    a.umax2 = umax2;
    const int nb = reduce_num_partials(nx, ny, nz);
    if (use_index32(nx, ny, nz))
        hipLaunchKernelGGL((umax2_kernel<T, unsigned>), dim3(nb), dim3(RED_TPB), 0, s, a);
    else
        hipLaunchKernelGGL((umax2_kernel<T, unsigned long long>), dim3(nb), dim3(RED_TPB), 0, s, a);
}

// ---------------------------------------------------------------- reference timestep inputs
// The (x, y) pair of cfl_cell for the current state (modes 1 and 2): |u|^2 and the Alfven speed
// squared |curl A|^2 / exp(lnrho), or (|u| + sqrt(cs^2 + vA^2))^2; maxima over interior cells.
template <typename T>
struct CflArgs {
    const T* f[NF];
    Coeffs<T> C;
    int nx, ny, nz, mx, my;
    int mode;
    double cs2;
    unsigned long long* out;  // pair
};

template <typename T, typename I>
__global__ void __launch_bounds__(RED_TPB) cfl_state_kernel(CflArgs<T> a)
{
    const I n = (I)a.nx * (I)a.ny * (I)a.nz;
    const I sy = (I)a.mx, sz = (I)a.mx * (I)a.my;
    double mx_ = 0.0, my_ = 0.0;
    for (I t = (I)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (I)gridDim.x * blockDim.x) {
        const I c = interior_index<I>(t, (I)a.nx, (I)a.ny, sy, sz);
        const I st[3] = {1, sy, sz};
        CellDerivs<T> d;
        d.lr = a.f[0][c];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            d.u[q] = a.f[1 + q][c];
            const T* f = a.f[4 + q];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const I s = st[ax];
                const T* kk = a.C.d1[ax];
                d.ga[q][ax] = kk[1] * (f[c + s] - f[c - s]) + kk[2] * (f[c + 2 * s] - f[c - 2 * s]) +
                              kk[3] * (f[c + 3 * s] - f[c - 3 * s]);
            }
        }
        double x, y;
        cfl_cell<T>(a.mode, a.cs2, d, x, y);
        mx_ = fmax(mx_, x);
        my_ = fmax(my_, y);
    }
    wave_max_to_global(mx_, a.out);
    if (a.mode == 1)
        wave_max_to_global(my_, a.out + 1);
}

template <typename T>
void launch_cfl_state(const T* const* fields, const Coeffs<T>& C, int nx, int ny, int nz, int mode,
                      double cs2, unsigned long long* out, hipStream_t s)
{
    CflArgs<T> a;
    for (int q = 0; q < NF; ++q)
        a.f[q] = fields[q];
    a.C = C;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.mx = nx + 2 * R;
    a.my = ny + 2 * R;
    a.mode = mode;
    a.cs2 = cs2;
    a.out = out;
    const long long n = (long long)nx * ny * nz;
    const int nb = (int)std::min<long long>((n + RED_TPB - 1) / RED_TPB, 4096);
    if (use_index32(nx, ny, nz))
        hipLaunchKernelGGL((cfl_state_kernel<T, unsigned>), dim3(nb), dim3(RED_TPB), 0, s, a);
    else
        hipLaunchKernelGGL((cfl_state_kernel<T, unsigned long long>), dim3(nb), dim3(RED_TPB), 0, s,
                           a);
}

template void launch_cfl_state<AcRealT>(const AcRealT* const*, const Coeffs<AcRealT>&, int, int, int,
                                         int, double, unsigned long long*, hipStream_t);

// ---------------------------------------------------------------- compensation records
template <typename E>
__global__ void zero_comp_field_kernel(E* first, size_t count)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < count;
         t += (size_t)gridDim.x * blockDim.x)
        first[t * NF] = E(0);
}

void launch_zero_comp_field(void* slab, size_t count, int kahan_bits, int q, hipStream_t s)
{
    size_t nb = (count + 255) / 256;
    if (nb > 65536)
        nb = 65536;
    char* first = (char*)slab + (size_t)q * (kahan_bits / 8);
    if (kahan_bits == 8)
        hipLaunchKernelGGL(zero_comp_field_kernel<int8_t>, dim3((unsigned)nb), dim3(256), 0, s,
                           (int8_t*)first, count);
    else if (kahan_bits == 16)
        hipLaunchKernelGGL(zero_comp_field_kernel<int16_t>, dim3((unsigned)nb), dim3(256), 0, s,
                           (int16_t*)first, count);
    else if (kahan_bits == 32)
        hipLaunchKernelGGL(zero_comp_field_kernel<float>, dim3((unsigned)nb), dim3(256), 0, s,
                           (float*)first, count);
}

// ---------------------------------------------------------------- helical forcing
// ref mhdsolver.ac:286-331 (helical_forcing, forcing) and :614-631 (added to u after the final
// update of substep 2). The geometry and the mode are evaluated in FP64, the force is rounded to
// AcReal and added to the stored u in AcReal, as the reference adds forcing_step to rk3_final.
template <typename T>
__global__ void __launch_bounds__(256) forcing_kernel(ForcingArgs<T> a)
{
    const int i = threadIdx.x + blockIdx.x * blockDim.x;
    const int j = threadIdx.y + blockIdx.y * blockDim.y;
    const int k = blockIdx.z;
    const bool inside = i < a.nx && j < a.ny && k < a.nz;
    const double dt = (double)solve_dt(a.s);
    const double t0 = a.time_acc ? a.t_base + (*a.time_acc - dt) : a.t_base;
    double u2 = 0.0;
    if (inside) {
        const long long c = (long long)(i + R) + (long long)a.mx * (j + R) + a.sz * (k + R);
        T u[3] = {a.uu[0][c], a.uu[1][c], a.uu[2][c]};
        if (t0 > a.switch_time) {
            // y = 2 pi xx / (ds ngrid), xx = (global index - nmin) ds
            const double g[3] = {(double)(a.gx0 + i), (double)(a.gy0 + j), (double)(a.gz0 + k)};
            const int ng[3] = {a.ngx, a.ngy, a.ngz};
            double kdy = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d)
                kdy += a.k[d] * (2.0 * M_PI * ((g[d] * a.ds[d]) / (a.ds[d] * (double)ng[d])));
            double sk, ck, sp, cp;
            sincos(kdy, &sk, &ck);
            sincos(a.phase, &sp, &cp);
            const double re = ck * cp - sk * sp;
            const double im = ck * sp + sk * cp;
            const double amp = sqrt(dt) * a.nn;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                u[d] = u[d] + (T)(amp * (a.fre[d] * re - a.fim[d] * im));
                a.uu[d][c] = u[d];
            }
        }
        u2 = (double)u[0] * (double)u[0] + (double)u[1] * (double)u[1] + (double)u[2] * (double)u[2];
    }
    if (a.umax2_next != nullptr)
        wave_max_to_global(u2, a.umax2_next);
}

template <typename T>
void launch_forcing(const ForcingArgs<T>& a, hipStream_t s)
{
    const dim3 tpb(64, 4, 1);
    const dim3 bpg((unsigned)((a.nx + 63) / 64), (unsigned)((a.ny + 3) / 4), (unsigned)a.nz);
    hipLaunchKernelGGL((forcing_kernel<T>), bpg, tpb, 0, s, a);
}

template void launch_forcing<AcRealT>(const ForcingArgs<AcRealT>&, hipStream_t);

// ---------------------------------------------------------------- fills
template <typename T>
__global__ void fill_kernel(T* dst, size_t n, T v)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (size_t)gridDim.x * blockDim.x)
        dst[t] = v;
}

template <typename T>
void launch_fill(T* dst, size_t n, T v, hipStream_t s)
{
    if (n == 0)
        return;
    size_t nb = (n + 255) / 256;
    if (nb > 65536)
        nb = 65536;
    hipLaunchKernelGGL(fill_kernel<T>, dim3((unsigned)nb), dim3(256), 0, s, dst, n, v);
}

void launch_zero_bytes(void* dst, size_t nbytes, hipStream_t s)
{
    // Word fill for the aligned body, byte fill for a tail (slabs are 256-byte aligned).
    const size_t words = nbytes / 4;
    launch_fill<unsigned>((unsigned*)dst, words, 0u, s);
    const size_t tail = nbytes - words * 4;
    if (tail)
        launch_fill<unsigned char>((unsigned char*)dst + words * 4, tail, (unsigned char)0, s);
}

template void launch_periodic_fill<AcRealT>(AcRealT* const*, int, int, int, int, hipStream_t);
template void launch_periodic_fill_box<AcRealT>(AcRealT* const*, int, int, int, int, const int*,
                                                  const int*, hipStream_t);
template void launch_reduce<AcRealT>(const AcRealT* const*, int, int, int, int, int, double*, int,
                                     double*, hipStream_t);
template void launch_diagnostics<AcRealT>(const AcRealT* const*, const Coeffs<AcRealT>&, int, int,
                                          int, double*, int, double*, hipStream_t);
template void launch_umax2<AcRealT>(const AcRealT* const*, int, int, int, unsigned long long*,
                                    hipStream_t);
template void launch_fill<AcRealT>(AcRealT*, size_t, AcRealT, hipStream_t);

}  // namespace acamd
