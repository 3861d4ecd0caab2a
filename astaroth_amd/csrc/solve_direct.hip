// solve_direct.hip -- one cell per thread fused RK3 substep, stencil points read straight from
// global memory (served by L1/L2). This is the simple, obviously-correct solve kernel: it is
// the parity anchor for the LDS-tiled kernel and the fallback for grids the tiled kernel does
// not cover. Semantics: ref acc-runtime/samples/mhd_modular/mhdsolver.ac:586-636 fused into one
// pass (ref twopass_solve_intermediate + twopass_solve_final).
#include "mhd_common.h"
#include "kernels.h"

namespace acamd {

template <typename T>
__device__ __forceinline__ T der1(const T* __restrict__ f, long long c, long long s, const T* k)
{
    return k[1] * (f[c + s] - f[c - s]) + k[2] * (f[c + 2 * s] - f[c - 2 * s]) +
           k[3] * (f[c + 3 * s] - f[c - 3 * s]);
}
// Second (and upwind) derivative. CF: the cancellation-free form of the precise 32-bit mode,
// sum_o k_o ((f_o - f_0) + (f_-o - f_0)), whose weights sum to zero like the sum form's, so the
// centre weight drops out; differences of neighbouring values are exact in FP32.
template <bool CF, typename T>
__device__ __forceinline__ T der2(const T* __restrict__ f, long long c, long long s, const T* k)
{
    if constexpr (CF) {
        const T f0 = f[c];
        return k[1] * ((f[c + s] - f0) + (f[c - s] - f0)) +
               k[2] * ((f[c + 2 * s] - f0) + (f[c - 2 * s] - f0)) +
               k[3] * ((f[c + 3 * s] - f0) + (f[c - 3 * s] - f0));
    }
    return k[0] * f[c] + k[1] * (f[c + s] + f[c - s]) + k[2] * (f[c + 2 * s] + f[c - 2 * s]) +
           k[3] * (f[c + 3 * s] + f[c - 3 * s]);
}
// Mixed derivative; CF pairs the corners as (f(o,o) - f(o,-o)) - (f(-o,o) - f(-o,-o)).
template <bool CF, typename T>
__device__ __forceinline__ T derm(const T* __restrict__ f, long long c, long long sa, long long sb,
                                  const T* k)
{
    T r = T(0);
#pragma unroll
    for (int o = 1; o <= 3; ++o) {
        if constexpr (CF)
            r += k[o] * ((f[c + o * sa + o * sb] - f[c + o * sa - o * sb]) -
                         (f[c - o * sa + o * sb] - f[c - o * sa - o * sb]));
        else
            r += k[o] * ((f[c + o * sa + o * sb] + f[c - o * sa - o * sb]) -
                         (f[c + o * sa - o * sb] + f[c - o * sa + o * sb]));
    }
    return r;
}

template <typename T, bool UPWD, bool ENTROPY, int KB>
__global__ void __launch_bounds__(256) solve_direct_kernel(SolveArgs<T> a)
{
    const int i = threadIdx.x + blockIdx.x * blockDim.x;
    const int j = threadIdx.y + blockIdx.y * blockDim.y;
    const int k = blockIdx.z;
    const bool active = i < a.nx && j < a.ny && k < a.nz;
    // the precise mode (32-bit compensation records, FP32)
    constexpr bool CF = sizeof(T) == 4 && KB == 32;
    double u2 = 0.0;  // max |u_new|^2 (the courant timestep mode)
    const T dt = solve_dt(a);
    if (active) {
        const long long sx = 1, sy = a.mx, sz = a.sz;
        const long long c = (long long)(i + R) + sy * (j + R) + sz * (k + R);
        const long long st[3] = {sx, sy, sz};
        const Coeffs<T>& C = a.C;

        CellDerivs<T> d;
        T grad[NF][3], lap[NF], upw[4][3];
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            const T* f = a.in[q];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax)
                grad[q][ax] = der1(f, c, st[ax], C.d1[ax]);
            lap[q] = der2<CF>(f, c, sx, C.d2[0]) + der2<CF>(f, c, sy, C.d2[1]) + der2<CF>(f, c, sz, C.d2[2]);
        }
        if (UPWD) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int ax = 0; ax < 3; ++ax)
                    upw[q][ax] = der2<CF>(a.in[q], c, st[ax], C.up[ax]);
        }
        T val[NF];
#pragma unroll
        for (int q = 0; q < NF; ++q)
            val[q] = a.in[q][c];

        d.lr = val[0];
        d.laplr = lap[0];
        d.s = val[7];
        d.laps = lap[7];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            d.glr[ax] = grad[0][ax];
            d.gs[ax] = grad[7][ax];
            d.u[ax] = val[1 + ax];
            d.lapu[ax] = lap[1 + ax];
            d.lapa[ax] = lap[4 + ax];
#pragma unroll
            for (int bx = 0; bx < 3; ++bx) {
                d.gu[ax][bx] = grad[1 + ax][bx];
                d.ga[ax][bx] = grad[4 + ax][bx];
            }
        }
        if (UPWD) {
            const T au[3] = {ac_abs(val[1]), ac_abs(val[2]), ac_abs(val[3])};
            d.upl = au[0] * upw[0][0] + au[1] * upw[0][1] + au[2] * upw[0][2];
#pragma unroll
            for (int q = 0; q < 3; ++q)
                d.upu[q] = upw[1 + q][0] * au[0] + upw[1 + q][1] * au[1] + upw[1 + q][2] * au[2];
        }
        else {
            d.upl = T(0);
            d.upu[0] = d.upu[1] = d.upu[2] = T(0);
        }
        // grad div: second derivatives on the diagonal, mixed derivatives off it
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const T* fx = a.in[1 + 3 * v];
            const T* fy = a.in[2 + 3 * v];
            const T* fz = a.in[3 + 3 * v];
            T* gd = v == 0 ? d.gdu : d.gda;
            gd[0] = der2<CF>(fx, c, sx, C.d2[0]) + derm<CF>(fy, c, sx, sy, C.mx[0]) +
                    derm<CF>(fz, c, sx, sz, C.mx[1]);
            gd[1] = derm<CF>(fx, c, sx, sy, C.mx[0]) + der2<CF>(fy, c, sy, C.d2[1]) +
                    derm<CF>(fz, c, sy, sz, C.mx[2]);
            gd[2] = derm<CF>(fx, c, sx, sz, C.mx[1]) + derm<CF>(fy, c, sy, sz, C.mx[2]) +
                    der2<CF>(fz, c, sz, C.d2[2]);
        }

        T rhs[NF];
        mhd_rhs<T, UPWD, ENTROPY, CF>(d, a.P, rhs);

        T fnew[NF];
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            if (q == 7 && !ENTROPY) {
                fnew[q] = val[q];
                a.out[q][c] = val[q];
                continue;
            }
            rk_update<T, KB>(a, q, c, val[q], rhs[q], dt, &fnew[q]);
            a.out[q][c] = fnew[q];
        }
        u2 = (double)fnew[1] * (double)fnew[1] + (double)fnew[2] * (double)fnew[2] +
             (double)fnew[3] * (double)fnew[3];
    }
    if (a.step == 2) {
        if (a.umax2_next != nullptr)
            wave_max_to_global(u2, a.umax2_next);
        if (a.time_acc != nullptr && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 &&
            threadIdx.x == 0 && threadIdx.y == 0)
            *a.time_acc += (double)dt;
    }
}

template <typename T, bool UPWD, bool ENTROPY, int KB>
static void launch_direct(const SolveArgs<T>& a, hipStream_t stream)
{
    const dim3 tpb(64, 4, 1);
    const dim3 bpg((a.nx + 63) / 64, (a.ny + 3) / 4, a.nz);
    hipLaunchKernelGGL((solve_direct_kernel<T, UPWD, ENTROPY, KB>), bpg, tpb, 0, stream, a);
}

template <typename T>
void launch_solve_direct(const SolveArgs<T>& a, bool upwd, bool entropy, int kb, hipStream_t s)
{
#define AC_DISPATCH_KB(U, E)                                                          \
    do {                                                                             \
        if (kb == 0) launch_direct<T, U, E, 0>(a, s);                                \
        else if (kb == 8) launch_direct<T, U, E, (sizeof(T) == 4 ? 8 : 0)>(a, s);    \
        else if (kb == 16) launch_direct<T, U, E, (sizeof(T) == 4 ? 16 : 0)>(a, s);  \
        else launch_direct<T, U, E, (sizeof(T) == 4 ? 32 : 0)>(a, s);                \
    } while (0)
    if (upwd && entropy) AC_DISPATCH_KB(true, true);
    else if (upwd) AC_DISPATCH_KB(true, false);
    else if (entropy) AC_DISPATCH_KB(false, true);
    else AC_DISPATCH_KB(false, false);
#undef AC_DISPATCH_KB
}

template void launch_solve_direct<AcRealT>(const SolveArgs<AcRealT>&, bool, bool, int, hipStream_t);

}  // namespace acamd
