// adjoint.hip -- tangent-linear and adjoint of one fused RK3 substep (no reference counterpart).
//
// Forward substep s (per field q, cell c):  w' = alpha_s w + dt R(f),  f' = f + beta_s w'.
// Tangent:  dw' = alpha_s dw + dt (J(f) df + dR/dnu dnu + dR/deta deta),  df' = df + beta_s dw'.
// Adjoint:  mu = lw' + beta_s lf',  lw = alpha_s mu,  lf = lf' + dt J(f)^T mu,
//           lnu += dt <mu, dR/dnu>,  leta += dt <mu, dR/deta>.
// J^T mu is computed in three stages: (A) per cell, the adjoint of every derivative quantity
// the right-hand side reads at a neighbour (NG = 47 fields, "derivative adjoints"), together
// with the pointwise parts: lw = alpha mu and lf += dt * (the adjoint of the cell's own value);
// (B) the periodic halo fill of the derivative adjoints, which is the transpose of the forward
// halo copy; (C) per output field, the transposed stencils applied to the derivative adjoints
// by a z-marching kernel, added to lf (first derivatives are antisymmetric, so their transpose
// is the negative; second, upwind and mixed ones are symmetric).
#include <cstdio>
#include <cstring>

#include "kernels.h"
#include "mhd_adjoint.h"

namespace acamd {


template <typename T>
struct AdjArgs {
    SolveArgs<T> s;        // f = s.in (halos filled), coefficients, parameters, alpha, beta
    T* lf;                 // adjoint of f, NF slabs of m^3, updated in place
    T* lw;                 // adjoint of w, updated in place
    T* g;                  // NG slabs of m^3 (derivative adjoints)
    double* partials;      // 2 per block (nu, eta)
    const T* df_in;        // tangent: df (halos filled)
    T* df_out;             // tangent: df'
    T* dw;                 // tangent: dw, updated in place
    T dnu, deta;
    long long count;       // m^3
};

template <typename T>
__device__ __forceinline__ T d1(const T* __restrict__ f, long long c, long long s, const T* k)
{
    return k[1] * (f[c + s] - f[c - s]) + k[2] * (f[c + 2 * s] - f[c - 2 * s]) +
           k[3] * (f[c + 3 * s] - f[c - 3 * s]);
}
template <typename T>
__device__ __forceinline__ T d2(const T* __restrict__ f, long long c, long long s, const T* k)
{
    return k[0] * f[c] + k[1] * (f[c + s] + f[c - s]) + k[2] * (f[c + 2 * s] + f[c - 2 * s]) +
           k[3] * (f[c + 3 * s] + f[c - 3 * s]);
}
template <typename T>
__device__ __forceinline__ T dm(const T* __restrict__ f, long long c, long long sa, long long sb,
                                const T* k)
{
    T r = T(0);
#pragma unroll
    for (int o = 1; o <= 3; ++o)
        r += k[o] * ((f[c + o * sa + o * sb] + f[c - o * sa - o * sb]) -
                     (f[c + o * sa - o * sb] + f[c - o * sa + o * sb]));
    return r;
}

// All derivative quantities of the eight fields at cell c (field q at base + q * count).
template <typename T, bool UPWD>
__device__ __forceinline__ void gather(const T* const* f, long long c, long long sy, long long sz,
                                       const Coeffs<T>& C, CellDerivs<T>& d, UpwindDirs<T>& up)
{
    const long long st[3] = {1, sy, sz};
    T val[NF], grad[NF][3], lap[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q) {
        val[q] = f[q][c];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
            grad[q][ax] = d1(f[q], c, st[ax], C.d1[ax]);
        lap[q] = d2(f[q], c, st[0], C.d2[0]) + d2(f[q], c, st[1], C.d2[1]) + d2(f[q], c, st[2], C.d2[2]);
    }
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
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        up.ul[ax] = UPWD ? d2(f[0], c, st[ax], C.up[ax]) : T(0);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            up.uu[i][ax] = UPWD ? d2(f[1 + i], c, st[ax], C.up[ax]) : T(0);
    }
#pragma unroll
    for (int v = 0; v < 2; ++v) {
        const T* fx = f[1 + 3 * v];
        const T* fy = f[2 + 3 * v];
        const T* fz = f[3 + 3 * v];
        T* gd = v == 0 ? d.gdu : d.gda;
        gd[0] = d2(fx, c, st[0], C.d2[0]) + dm(fy, c, st[0], st[1], C.mx[0]) + dm(fz, c, st[0], st[2], C.mx[1]);
        gd[1] = dm(fx, c, st[0], st[1], C.mx[0]) + d2(fy, c, st[1], C.d2[1]) + dm(fz, c, st[1], st[2], C.mx[2]);
        gd[2] = dm(fx, c, st[0], st[2], C.mx[1]) + dm(fy, c, st[1], st[2], C.mx[2]) + d2(fz, c, st[2], C.d2[2]);
    }
    combine_upwind(d, up);
}

template <typename T>
__device__ __forceinline__ void cell_index(const SolveArgs<T>& a, bool& active, long long& c)
{
    const int i = threadIdx.x + blockIdx.x * blockDim.x;
    const int j = threadIdx.y + blockIdx.y * blockDim.y;
    const int k = blockIdx.z;
    active = i < a.nx && j < a.ny && k < a.nz;
    c = (long long)(i + R) + (long long)a.mx * (j + R) + a.sz * (k + R);
}

// ---------------------------------------------------------------- tangent
template <typename T, bool UPWD, bool ENTROPY>
__global__ void __launch_bounds__(256) tangent_kernel(AdjArgs<T> A)
{
    const SolveArgs<T>& a = A.s;
    bool active;
    long long c;
    cell_index(a, active, c);
    if (!active)
        return;
    CellDerivs<T> d, dd;
    UpwindDirs<T> up, dup;
    gather<T, UPWD>(a.in, c, a.mx, a.sz, a.C, d, up);
    const T* df[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q)
        df[q] = A.df_in + (long long)q * A.count;
    gather<T, UPWD>(df, c, a.mx, a.sz, a.C, dd, dup);
    T rdot[NF];
    mhd_rhs_jvp<T, UPWD, ENTROPY>(d, up, dd, dup, a.P, A.dnu, A.deta, rdot);
#pragma unroll
    for (int q = 0; q < NF; ++q) {
        const long long o = (long long)q * A.count + c;
        if (q == 7 && !ENTROPY) {
            A.df_out[o] = df[q][c];
            continue;
        }
        T dwn = a.dt * rdot[q];
        if (a.step > 0)
            dwn = a.alpha * A.dw[o] + dwn;
        if (a.step < 2)
            A.dw[o] = dwn;
        A.df_out[o] = df[q][c] + a.beta * dwn;
    }
}

// ---------------------------------------------------------------- adjoint (A): local part
template <typename T, bool UPWD, bool ENTROPY>
__global__ void __launch_bounds__(256) adjoint_local_kernel(AdjArgs<T> A)
{
    __shared__ double sh[2][4];
    const SolveArgs<T>& a = A.s;
    bool active;
    long long c;
    cell_index(a, active, c);
    double pnu = 0.0, peta = 0.0;
    if (active) {
        CellDerivs<T> d, ad;
        UpwindDirs<T> up, aup;
        gather<T, UPWD>(a.in, c, a.mx, a.sz, a.C, d, up);
        T mu[NF];
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            const long long o = (long long)q * A.count + c;
            mu[q] = A.lw[o] + a.beta * A.lf[o];
        }
        if (!ENTROPY)
            mu[7] = T(0);
        T anu, aeta;
        mhd_rhs_vjp<T, UPWD, ENTROPY>(d, up, mu, a.P, ad, aup, anu, aeta);
        pnu = (double)a.dt * (double)anu;
        peta = (double)a.dt * (double)aeta;
        T* g = A.g;
        const long long n = A.count;
        put_derivative_adjoints<UPWD>(ad, aup, [&](int k, T v) { g[(long long)k * n + c] = v; });
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            const long long o = (long long)q * n + c;
            const T lfo = A.lf[o];
            A.lw[o] = a.alpha * mu[q];
            if (q < 4 || (q == 7 && ENTROPY))
                A.lf[o] = lfo + a.dt * value_adjoint(ad, q);
        }
    }
    // block sums of the parameter sensitivities (FP64)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        pnu += __shfl_xor(pnu, off, 64);
        peta += __shfl_xor(peta, off, 64);
    }
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    if ((tid & 63) == 0) {
        sh[0][tid >> 6] = pnu;
        sh[1][tid >> 6] = peta;
    }
    __syncthreads();
    if (tid == 0) {
        const int nw = (blockDim.x * blockDim.y + 63) >> 6;
        double s0 = 0.0, s1 = 0.0;
        for (int w = 0; w < nw; ++w) {
            s0 += sh[0][w];
            s1 += sh[1][w];
        }
        const long long b = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
        A.partials[2 * b + 0] = s0;
        A.partials[2 * b + 1] = s1;
    }
}

// Deterministic two-stage sum of the (nu, eta) block partials: stage 1 reduces contiguous
// ranges of partials per workgroup, stage 2 (one workgroup) adds the range sums into out.
constexpr int PARAM_RANGES = 256;

__device__ __forceinline__ void block_sum2(double& s0, double& s1)
{
    __shared__ double sh[2][4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off, 64);
        s1 += __shfl_xor(s1, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = s0;
        sh[1][threadIdx.x >> 6] = s1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0.0, t1 = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            t0 += sh[0][w];
            t1 += sh[1][w];
        }
        s0 = t0;
        s1 = t1;
    }
}

__global__ void __launch_bounds__(256) adjoint_param_stage1(const double* partials, long long nb,
                                                            double* ranges)
{
    const long long per = (nb + PARAM_RANGES - 1) / PARAM_RANGES;
    const long long b0 = per * blockIdx.x;
    const long long b1 = b0 + per < nb ? b0 + per : nb;
    double s0 = 0.0, s1 = 0.0;
    for (long long t = b0 + threadIdx.x; t < b1; t += blockDim.x) {
        s0 += partials[2 * t];
        s1 += partials[2 * t + 1];
    }
    block_sum2(s0, s1);
    if (threadIdx.x == 0) {
        ranges[2 * blockIdx.x] = s0;
        ranges[2 * blockIdx.x + 1] = s1;
    }
}

__global__ void __launch_bounds__(256) adjoint_param_stage2(const double* ranges, double* out)
{
    double s0 = 0.0, s1 = 0.0;
    for (int t = threadIdx.x; t < PARAM_RANGES; t += blockDim.x) {
        s0 += ranges[2 * t];
        s1 += ranges[2 * t + 1];
    }
    block_sum2(s0, s1);
    if (threadIdx.x == 0) {
        out[0] += s0;
        out[1] += s1;
    }
}

// ---------------------------------------------------------------- adjoint (C), z-marching form
// The same transposed stencils, organised so that every derivative-adjoint value is read from
// memory once per use pattern: a thread owns an (x, y) column and walks a chunk of planes. Each
// output field q has a chain acc[q][k] of partial sums for planes p - 3 + k (k = 0..6), where p
// is the plane being read. Reading plane p pushes
//   * its pointwise and in-plane (x / y) contributions into acc[q][3] (plane p itself);
//   * its z-stencil contributions into acc[q][3 -+ i]: a value a(p) under the transposed first
//     derivative lands on plane p - i with -d1_i and on p + i with +d1_i; under second
//     derivative, upwind and D2z of grad(div) with the symmetric weights;
//   * the xz / yz mixed parts: D_i G(p) = G(x+i, p) - G(x-i, p) lands on p - i with +c_i and on
//     p + i with -c_i.
// Plane p - 3 is then complete (acc[q][0]) and its adjoint update is written.
template <typename T, bool UPWD, int Q>
__global__ void __launch_bounds__(256) adjoint_transpose_z_kernel(AdjArgs<T> A, int zchunk)
{
    const SolveArgs<T>& a = A.s;
    const int i = threadIdx.x + blockIdx.x * blockDim.x;
    const int j = threadIdx.y + blockIdx.y * blockDim.y;
    const int z0 = blockIdx.z * zchunk;
    if (i >= a.nx || j >= a.ny)
        return;
    const int sy = a.mx;
    const long long sz = a.sz;
    const Coeffs<T>& C = a.C;
    const unsigned colb0 = (unsigned)(((i + R) + sy * (j + R)) * (int)sizeof(T));
    T acc[7];
#pragma unroll
    for (int k = 0; k < 7; ++k)
        acc[k] = T(0);

    const int np = zchunk + 2 * R;
    for (int it = 0; it < np; ++it) {
        const int p = z0 - R + it;  // interior plane index being read (array plane p + R)
        // uniform bases recomputed per plane from an opaque field stride; 32-bit lane offsets
        const long long nb = opaque_s64(A.count * (long long)sizeof(T));
        const unsigned colb = opaque(colb0);
        const char* const pb = (const char*)A.g + sz * (long long)(p + R) * (long long)sizeof(T);
        // value of derivative-adjoint field k at offset (dx, dy) from this column in plane p;
        // the signed neighbour offset is added to the column's byte offset in 32-bit unsigned
        // arithmetic first (the sum is a valid non-negative offset), then to the base pointer
        auto ld = [&](int k, int dx, int dy) -> T {
            const unsigned off = colb + (unsigned)((dx + dy * sy) * (int)sizeof(T));
            return *(const T*)(pb + k * nb + off);
        };
        auto d1x = [&](int k) {
            return C.d1[0][1] * (ld(k, 1, 0) - ld(k, -1, 0)) + C.d1[0][2] * (ld(k, 2, 0) - ld(k, -2, 0)) +
                   C.d1[0][3] * (ld(k, 3, 0) - ld(k, -3, 0));
        };
        auto d1y = [&](int k) {
            return C.d1[1][1] * (ld(k, 0, 1) - ld(k, 0, -1)) + C.d1[1][2] * (ld(k, 0, 2) - ld(k, 0, -2)) +
                   C.d1[1][3] * (ld(k, 0, 3) - ld(k, 0, -3));
        };
        // symmetric stencil along x (ax = 0) or y (ax = 1), centre weight included
        auto sym = [&](int k, int ax, const T* w) {
            const int ex = ax == 0, ey = ax == 1;
            return w[0] * ld(k, 0, 0) + w[1] * (ld(k, ex, ey) + ld(k, -ex, -ey)) +
                   w[2] * (ld(k, 2 * ex, 2 * ey) + ld(k, -2 * ex, -2 * ey)) +
                   w[3] * (ld(k, 3 * ex, 3 * ey) + ld(k, -3 * ex, -3 * ey));
        };
        auto mixed_xy = [&](int k) {
            T r = T(0);
#pragma unroll
            for (int o = 1; o <= 3; ++o)
                r += C.mx[0][o] * ((ld(k, o, o) + ld(k, -o, -o)) - (ld(k, o, -o) + ld(k, -o, o)));
            return r;
        };
        // transposed first derivative along z of a column value: -d1_o on p - o, +d1_o on p + o
        auto push_anti = [&](T v) {
#pragma unroll
            for (int o = 1; o <= 3; ++o) {
                acc[3 - o] -= C.d1[2][o] * v;
                acc[3 + o] += C.d1[2][o] * v;
            }
        };
        auto push_sym = [&](T v, const T* w) {
            acc[3] += w[0] * v;
#pragma unroll
            for (int o = 1; o <= 3; ++o) {
                acc[3 - o] += w[o] * v;
                acc[3 + o] += w[o] * v;
            }
        };
        // xz (ax = 0) or yz (ax = 1) mixed part of field k: D_o G(p) on p - o with +c_o, on
        // p + o with -c_o
        auto push_mixed = [&](int k, int ax) {
            const int ex = ax == 0, ey = ax == 1;
            const T* w = C.mx[ax == 0 ? 1 : 2];
#pragma unroll
            for (int o = 1; o <= 3; ++o) {
                const T D = w[o] * (ld(k, o * ex, o * ey) - ld(k, -o * ex, -o * ey));
                acc[3 - o] += D;
                acc[3 + o] -= D;
            }
        };

        // own stencils: first derivatives in x / y (absent for d_i A_i), the Laplacian, the
        // upwind operators (lnrho and u); the z parts pushed into the chain
        {
            constexpr int kx = g_index(Q, ROLE_DX), ky = g_index(Q, ROLE_DY);
            constexpr int kz = g_index(Q, ROLE_DZ), kl = g_index(Q, ROLE_LAP);
            T r = sym(kl, 0, C.d2[0]) + sym(kl, 1, C.d2[1]);
            if constexpr (kx >= 0)
                r -= d1x(kx);
            if constexpr (ky >= 0)
                r -= d1y(ky);
            if constexpr (UPWD && Q < 4)
                r += sym(g_index(Q, ROLE_UPX), 0, C.up[0]) + sym(g_index(Q, ROLE_UPY), 1, C.up[1]);
            acc[3] += r;
            if constexpr (kz >= 0)
                push_anti(ld(kz, 0, 0));
            push_sym(ld(kl, 0, 0), C.d2[2]);
            if constexpr (UPWD && Q < 4)
                push_sym(ld(g_index(Q, ROLE_UPZ), 0, 0), C.up[2]);
        }
        // grad(div v) parts of the vector components (u: Q = 1..3 from G_GDU, A: 4..6 from G_GDA)
        if (Q >= 1 && Q <= 6) {
            const int gb = Q <= 3 ? G_GDU : G_GDA;
            const int comp = (Q - 1) % 3;
            if (comp == 0) {
                acc[3] += sym(gb + 0, 0, C.d2[0]) + mixed_xy(gb + 1);
                push_mixed(gb + 2, 0);  // d_xz of G2
            }
            else if (comp == 1) {
                acc[3] += mixed_xy(gb + 0) + sym(gb + 1, 1, C.d2[1]);
                push_mixed(gb + 2, 1);  // d_yz of G2
            }
            else {
                push_mixed(gb + 0, 0);  // d_xz of G0
                push_mixed(gb + 1, 1);  // d_yz of G1
                push_sym(ld(gb + 2, 0, 0), C.d2[2]);
            }
        }

        // ---- plane p - 3 is complete: its adjoint update
        if (it >= 2 * R) {
            const long long ob = sz * (long long)p * (long long)sizeof(T);  // array plane p - 3 + R
            T* lfp = (T*)((char*)A.lf + Q * nb + ob + colb);
            *lfp = *lfp + a.dt * acc[0];
        }
#pragma unroll
        for (int k = 0; k < 6; ++k)
            acc[k] = acc[k + 1];
        acc[6] = T(0);
    }
}

// ---------------------------------------------------------------- adjoint (C), LDS-staged form
// The same per-output-field transposes on 64 x 16, 64 x 8 or 64 x 4 column blocks: the
// derivative-adjoint fields output field Q reads with in-plane (x / y) neighbours are DMA'd
// plane by plane into LDS (with their radius-3 halo, double-buffered, as the forward solve
// stages its planes), fields read only at the column come straight from memory, and the z parts
// use the same push chain.

// Fields of output Q that are read with in-plane neighbours ("staged"), in slot order: its own
// x / y first derivatives (those that exist), Laplacian and, for lnrho and u with upwinding, the
// x / y upwind operators; then the grad(div) fields G0, G1[, G2] of the vector (G2 only for the
// x and y components).
__host__ __device__ constexpr int own_staged_role(int Q, bool upwd, int slot)
{
    int n = 0;
    for (int role : {ROLE_DX, ROLE_DY, ROLE_LAP, ROLE_UPX, ROLE_UPY}) {
        if ((role == ROLE_UPX || role == ROLE_UPY) && !(upwd && Q < 4))
            continue;
        if (g_index(Q, role) < 0)
            continue;
        if (n == slot)
            return role;
        ++n;
    }
    return -1;
}
__host__ __device__ constexpr int own_staged_count(int Q, bool upwd)
{
    int n = 0;
    while (own_staged_role(Q, upwd, n) >= 0)
        ++n;
    return n;
}
// Slot of an own role among the staged fields of Q, or -1.
__host__ __device__ constexpr int staged_slot(int Q, bool upwd, int role)
{
    for (int sl = 0; sl < own_staged_count(Q, upwd); ++sl)
        if (own_staged_role(Q, upwd, sl) == role)
            return sl;
    return -1;
}
__host__ __device__ constexpr int staged_count(int Q, bool upwd)
{
    const int gd = (Q >= 1 && Q <= 6) ? ((Q - 1) % 3 == 2 ? 2 : 3) : 0;
    return own_staged_count(Q, upwd) + gd;
}
__host__ __device__ constexpr int staged_field(int Q, bool upwd, int slot)
{
    const int no = own_staged_count(Q, upwd);
    if (slot < no)
        return g_index(Q, own_staged_role(Q, upwd, slot));
    return (Q <= 3 ? G_GDU : G_GDA) + (slot - no);
}

// Fields of output Q read only at the column itself, in slot order: its own z first derivative
// (where it exists) and z upwind operator (lnrho and u with upwinding); then G2 of grad(div) for
// the z components.
__host__ __device__ constexpr int own_column_count(int Q, bool upwd)
{
    return (g_index(Q, ROLE_DZ) >= 0 ? 1 : 0) + ((upwd && Q < 4) ? 1 : 0);
}
__host__ __device__ constexpr int column_count(int Q, bool upwd)
{
    return own_column_count(Q, upwd) + ((Q == 3 || Q == 6) ? 1 : 0);
}
__host__ __device__ constexpr int column_field(int Q, bool upwd, int slot)
{
    const bool dz = g_index(Q, ROLE_DZ) >= 0;
    if (dz && slot == 0)
        return g_index(Q, ROLE_DZ);
    if (slot < own_column_count(Q, upwd))
        return g_index(Q, ROLE_UPZ);
    return (Q <= 3 ? G_GDU : G_GDA) + 2;
}

// TY = 16 (FP32) or 8 where ny allows (1024- or 512-thread workgroups: the staged in-plane halo
// drops from 10/4 to 22/16 or 14/8 rows per output row, and FP64 runs two waves per SIMD
// instead of one), else 4.
template <typename T, bool UPWD, int Q, int TY>
__global__ void __launch_bounds__(64 * TY) adjoint_transpose_lds_kernel(AdjArgs<T> A, int zchunk,
                                                                        int ntx, int nty)
{
    constexpr int TX = 64;
    constexpr int LX = TX + 2 * R, LY = TY + 2 * R;
    constexpr int LP = LX * LY;
    constexpr int NT = TX * TY, NW = NT / 64;
    constexpr int WPE = (int)sizeof(T) / 4;
    constexpr int LPW = LP * WPE;
    constexpr int NCHUNK = (LPW + 63) / 64;
    constexpr int NR = (NCHUNK + NW - 1) / NW;
    constexpr int LPS = NR * NW * 64 / WPE;  // padded field-plane stride in LDS (elements)
    constexpr int NS = staged_count(Q, UPWD);
    constexpr int NC = column_count(Q, UPWD);
    __shared__ T lds[2][NS * LPS];

    const SolveArgs<T>& a = A.s;
    // XCD-aware numbering: consecutive hardware block ids go to different XCDs, so each XCD is
    // handed a contiguous range of tiles (whose y neighbours then share its L2)
    const int nblocks = (int)gridDim.x;
    int bid = (int)blockIdx.x;
    if ((nblocks & 7) == 0)
        bid = (bid & 7) * (nblocks >> 3) + (bid >> 3);
    const int tile_x = bid % ntx;
    const int rest = bid / ntx;
    const int tile_y = rest % nty;
    const int chunk = rest / nty;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = tx + TX * ty;
    // wave-uniform: a scalar register in FP32 (scalar M0 arithmetic for the DMA); FP64 keeps
    // the vector form, as the forward kernels do
    const int wave = sizeof(T) == 4 ? __builtin_amdgcn_readfirstlane(tid >> 6) : tid >> 6;
    const int lane = tid & 63;
    const int x0 = tile_x * TX, y0 = tile_y * TY, z0 = chunk * zchunk;
    const int sy = a.mx;
    const long long sz = a.sz;
    const Coeffs<T>& C = a.C;

    // DMA: every lane moves one dword per chunk; padding lanes re-read dword 0 of the plane.
    // The x / y halo of a tile is read from the periodic images inside the interior, so the
    // derivative adjoints need no halo fill.
    unsigned g_off[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        int d = (wave + r * NW) * 64 + lane;
        if (d >= LPW)
            d = 0;
        const int e = d / WPE, wd = d - e * WPE;
        const int ly = e / LX, lx = e - ly * LX;
        int gx = x0 + lx - R, gy = y0 + ly - R;
        gx += gx < 0 ? a.nx : (gx >= a.nx ? -a.nx : 0);
        gy += gy < 0 ? a.ny : (gy >= a.ny ? -a.ny : 0);
        g_off[r] = (unsigned)(((gx + R) + sy * (gy + R)) * (int)sizeof(T) + 4 * wd);
    }
    // byte offset of interior plane p (z wrapped periodically) within a field
    auto plane_b = [&](int p) {
        p += p < 0 ? a.nz : (p >= a.nz ? -a.nz : 0);
        return sz * (long long)(p + R) * (long long)sizeof(T);
    };
    auto dma_plane = [&](int p, int b, long long nb) {
        const char* pb = (const char*)A.g + plane_b(p);
#pragma unroll
        for (int sl = 0; sl < NS; ++sl) {
            const char* src = pb + staged_field(Q, UPWD, sl) * nb;
            char* dst = (char*)&lds[b][sl * LPS];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const int j = wave + r * NW;
                __builtin_amdgcn_global_load_lds((const void*)(src + opaque(g_off[r])),
                                                 (lds_ptr_t)(dst + 256 * j), 4, 0, 0);
            }
        }
    };

    const int ctr = (ty + R) * LX + (tx + R);
    const unsigned colb0 = (unsigned)(((x0 + tx + R) + sy * (y0 + ty + R)) * (int)sizeof(T));
    T acc[7];
#pragma unroll
    for (int k = 0; k < 7; ++k)
        acc[k] = T(0);
    // column values of the plane being processed, and of the next one (loaded a plane ahead)
    T col[NC], coln[NC];
    // lambda_f of the plane completed in this step (lfc) and in the next one (loaded a plane
    // ahead); the local stage has already added its pointwise part
    T lfc = T(0), lfn = T(0);
    auto load_cols = [&](int p, long long nb, unsigned colb, T* dst) {
        const char* pb = (const char*)A.g + plane_b(p);
#pragma unroll
        for (int c = 0; c < NC; ++c)
            dst[c] = *(const T*)(pb + column_field(Q, UPWD, c) * nb + colb);
    };
    // the output plane of step it is p - 3 = z0 - 2R + it (array plane z0 - R + it)
    auto out_off = [&](int it, long long nb, unsigned colb) {
        return Q * nb + sz * (long long)(z0 - R + it) * (long long)sizeof(T) + colb;
    };

    {
        const long long nb = A.count * (long long)sizeof(T);
        dma_plane(z0 - R, 0, nb);
        load_cols(z0 - R, nb, colb0, col);
    }
    plane_barrier();
    const int np = zchunk + 2 * R;
    for (int it = 0; it < np; ++it) {
        const int p = z0 - R + it;
        const int cur = it & 1;
        const long long nb = opaque_s64(A.count * (long long)sizeof(T));
        const unsigned colb = opaque(colb0);
        if (it + 1 < np) {
            dma_plane(p + 1, cur ^ 1, nb);
            load_cols(p + 1, nb, colb, coln);
            if (it + 1 >= 2 * R) {
                const long long ob = out_off(it + 1, nb, colb);
                lfn = *(const T*)((const char*)A.lf + ob);
            }
        }
        auto lds_at = [&](int sl, int dx, int dy) -> T {
            return lds[cur][sl * LPS + ctr + dy * LX + dx];
        };
        auto d1x = [&](int sl) {
            return C.d1[0][1] * (lds_at(sl, 1, 0) - lds_at(sl, -1, 0)) +
                   C.d1[0][2] * (lds_at(sl, 2, 0) - lds_at(sl, -2, 0)) +
                   C.d1[0][3] * (lds_at(sl, 3, 0) - lds_at(sl, -3, 0));
        };
        auto d1y = [&](int sl) {
            return C.d1[1][1] * (lds_at(sl, 0, 1) - lds_at(sl, 0, -1)) +
                   C.d1[1][2] * (lds_at(sl, 0, 2) - lds_at(sl, 0, -2)) +
                   C.d1[1][3] * (lds_at(sl, 0, 3) - lds_at(sl, 0, -3));
        };
        auto sym = [&](int sl, int ax, const T* w) {
            const int ex = ax == 0, ey = ax == 1;
            return w[0] * lds_at(sl, 0, 0) + w[1] * (lds_at(sl, ex, ey) + lds_at(sl, -ex, -ey)) +
                   w[2] * (lds_at(sl, 2 * ex, 2 * ey) + lds_at(sl, -2 * ex, -2 * ey)) +
                   w[3] * (lds_at(sl, 3 * ex, 3 * ey) + lds_at(sl, -3 * ex, -3 * ey));
        };
        auto mixed_xy = [&](int sl) {
            T r = T(0);
#pragma unroll
            for (int o = 1; o <= 3; ++o)
                r += C.mx[0][o] * ((lds_at(sl, o, o) + lds_at(sl, -o, -o)) -
                                   (lds_at(sl, o, -o) + lds_at(sl, -o, o)));
            return r;
        };
        auto push_anti = [&](T v) {
#pragma unroll
            for (int o = 1; o <= 3; ++o) {
                acc[3 - o] -= C.d1[2][o] * v;
                acc[3 + o] += C.d1[2][o] * v;
            }
        };
        auto push_sym = [&](T v, const T* w) {
            acc[3] += w[0] * v;
#pragma unroll
            for (int o = 1; o <= 3; ++o) {
                acc[3 - o] += w[o] * v;
                acc[3 + o] += w[o] * v;
            }
        };
        auto push_mixed = [&](int sl, int ax) {
            const int ex = ax == 0, ey = ax == 1;
            const T* w = C.mx[ax == 0 ? 1 : 2];
#pragma unroll
            for (int o = 1; o <= 3; ++o) {
                const T D = w[o] * (lds_at(sl, o * ex, o * ey) - lds_at(sl, -o * ex, -o * ey));
                acc[3 - o] += D;
                acc[3 + o] -= D;
            }
        };

        // own stencils (staged and column slots as listed above)
        {
            constexpr int sx = staged_slot(Q, UPWD, ROLE_DX), sy_ = staged_slot(Q, UPWD, ROLE_DY);
            constexpr int sl = staged_slot(Q, UPWD, ROLE_LAP);
            constexpr int sux = staged_slot(Q, UPWD, ROLE_UPX), suy = staged_slot(Q, UPWD, ROLE_UPY);
            constexpr bool has_dz = g_index(Q, ROLE_DZ) >= 0;
            constexpr int cuz = has_dz ? 1 : 0;  // column slot of the z upwind operator
            T r = sym(sl, 0, C.d2[0]) + sym(sl, 1, C.d2[1]);
            if constexpr (sx >= 0)
                r -= d1x(sx);
            if constexpr (sy_ >= 0)
                r -= d1y(sy_);
            if constexpr (sux >= 0)
                r += sym(sux, 0, C.up[0]) + sym(suy, 1, C.up[1]);
            acc[3] += r;
            if constexpr (has_dz)
                push_anti(col[0]);
            push_sym(lds_at(sl, 0, 0), C.d2[2]);
            if constexpr (sux >= 0)
                push_sym(col[cuz], C.up[2]);
        }
        // grad(div v) parts (staged slots from the own-stencil count: G0, G1[, G2])
        if (Q >= 1 && Q <= 6) {
            constexpr int no = own_staged_count(Q, UPWD);
            const int comp = (Q - 1) % 3;
            if (comp == 0) {
                acc[3] += sym(no + 0, 0, C.d2[0]) + mixed_xy(no + 1);
                push_mixed(no + 2, 0);  // d_xz of G2
            }
            else if (comp == 1) {
                acc[3] += mixed_xy(no + 0) + sym(no + 1, 1, C.d2[1]);
                push_mixed(no + 2, 1);  // d_yz of G2
            }
            else {
                push_mixed(no + 0, 0);  // d_xz of G0
                push_mixed(no + 1, 1);  // d_yz of G1
                push_sym(col[NC - 1], C.d2[2]);
            }
        }

        if (it >= 2 * R) {
            const long long ob = out_off(it, nb, colb);
            *(T*)((char*)A.lf + ob) = lfc + a.dt * acc[0];
        }
#pragma unroll
        for (int k = 0; k < 6; ++k)
            acc[k] = acc[k + 1];
        acc[6] = T(0);
#pragma unroll
        for (int c = 0; c < NC; ++c)
            col[c] = coln[c];
        lfc = lfn;
        // plane p + 1 has landed and every wave is done with this one
        plane_barrier();
    }
}

// ---------------------------------------------------------------- launchers
template <typename T>
static dim3 grid_of(const SolveArgs<T>& a)
{
    return dim3((a.nx + 63) / 64, (a.ny + 3) / 4, a.nz);
}

template <typename T>
void launch_tangent(const AdjArgs<T>& A, bool upwd, bool entropy, hipStream_t s)
{
    const dim3 tpb(64, 4, 1), bpg = grid_of(A.s);
    if (upwd && entropy) hipLaunchKernelGGL((tangent_kernel<T, true, true>), bpg, tpb, 0, s, A);
    else if (upwd) hipLaunchKernelGGL((tangent_kernel<T, true, false>), bpg, tpb, 0, s, A);
    else if (entropy) hipLaunchKernelGGL((tangent_kernel<T, false, true>), bpg, tpb, 0, s, A);
    else hipLaunchKernelGGL((tangent_kernel<T, false, false>), bpg, tpb, 0, s, A);
}

long long adjoint_num_blocks(int nx, int ny, int nz)
{
    return (long long)((nx + 63) / 64) * ((ny + 3) / 4) * nz;
}

template <typename T>
void launch_adjoint(const AdjArgs<T>& A, bool upwd, bool entropy, double* param_grad,
                    hipStream_t s)
{
    const dim3 tpb(64, 4, 1), bpg = grid_of(A.s);
#define AC_ADJ(K)                                                                                  \
    do {                                                                                           \
        if (upwd && entropy) hipLaunchKernelGGL((K<T, true, true>), bpg, tpb, 0, s, A);            \
        else if (upwd) hipLaunchKernelGGL((K<T, true, false>), bpg, tpb, 0, s, A);                 \
        else if (entropy) hipLaunchKernelGGL((K<T, false, true>), bpg, tpb, 0, s, A);              \
        else hipLaunchKernelGGL((K<T, false, false>), bpg, tpb, 0, s, A);                          \
    } while (0)
    // (A) derivative adjoints: the tiled kernel where the grid fits its 64 x 4 tiles (it has at
    // most as many blocks, so its partial sums fit the same workspace), else one cell per thread
    long long nb = 0;
    const AdjLocal<T> xl{A.g, A.lf, A.lw, A.partials};
    if (!launch_adjoint_local_tiled<T>(A.s, xl, upwd, entropy, s, &nb)) {
        AC_ADJ(adjoint_local_kernel);
        nb = (long long)bpg.x * bpg.y * bpg.z;
    }
    double* ranges = A.partials + 2 * nb;
    hipLaunchKernelGGL(adjoint_param_stage1, dim3(PARAM_RANGES), dim3(256), 0, s,
                       (const double*)A.partials, nb, ranges);
    hipLaunchKernelGGL(adjoint_param_stage2, dim3(1), dim3(256), 0, s, (const double*)ranges,
                       param_grad);
    // (C) transposed stencils, z-marching over chunks of planes: the LDS-staged form where the
    // grid fits its 64 x 4 column blocks (it reads the periodic images itself), else the plain
    // form after (B) a periodic fill of the derivative adjoints
    const bool staged = A.s.nx % 64 == 0 && A.s.ny % 4 == 0 && A.s.nz >= R;
    for (int k0 = 0; !staged && k0 < NG; k0 += NF) {
        T* ptrs[NF];
        const int cnt = NG - k0 < NF ? NG - k0 : NF;
        for (int q = 0; q < cnt; ++q)
            ptrs[q] = A.g + (long long)(k0 + q) * A.count;
        launch_periodic_fill<T>(ptrs, cnt, A.s.nx, A.s.ny, A.s.nz, s);
    }
    // the transposes' column-block height: 16 (FP32) or 8 where ny allows, else 4
    constexpr int TYB = sizeof(T) == 4 ? 16 : 8;
    const int tyc = A.s.ny % TYB == 0 ? TYB : (A.s.ny % 8 == 0 ? 8 : 4);
    const int ntx = A.s.nx / 64, nty = A.s.ny / tyc;
    // z-chunks: 64 planes where that still gives the staged form at least 256 workgroups (one
    // per CU; 256^3: 9.20 -> 9.03 ms per adjoint step against 32-plane chunks), else the largest
    // of 32, 16 and 8 below nz
    int zc = A.s.nz;
    if (staged && A.s.nz % 64 == 0 && A.s.nz > 64 && (long long)ntx * nty * (A.s.nz / 64) >= 256) {
        zc = 64;
    }
    else {
        for (int cand : {32, 16, 8})
            if (A.s.nz % cand == 0 && A.s.nz > cand) {
                zc = cand;
                break;
            }
    }
    const dim3 bpz(bpg.x, bpg.y, (unsigned)(A.s.nz / zc));
    const dim3 bps((unsigned)(ntx * nty * (A.s.nz / zc)), 1, 1), tpbb(64, TYB, 1), tpb8(64, 8, 1);
#define AC_ADJ_Q(Q)                                                                                \
    do {                                                                                           \
        if (staged && tyc == TYB) {                                                                \
            if (upwd) hipLaunchKernelGGL((adjoint_transpose_lds_kernel<T, true, Q, TYB>), bps, tpbb, 0, s, A, zc, ntx, nty); \
            else hipLaunchKernelGGL((adjoint_transpose_lds_kernel<T, false, Q, TYB>), bps, tpbb, 0, s, A, zc, ntx, nty);    \
        }                                                                                          \
        else if (staged && tyc == 8) {                                                             \
            if (upwd) hipLaunchKernelGGL((adjoint_transpose_lds_kernel<T, true, Q, 8>), bps, tpb8, 0, s, A, zc, ntx, nty); \
            else hipLaunchKernelGGL((adjoint_transpose_lds_kernel<T, false, Q, 8>), bps, tpb8, 0, s, A, zc, ntx, nty);    \
        }                                                                                          \
        else if (staged) {                                                                         \
            if (upwd) hipLaunchKernelGGL((adjoint_transpose_lds_kernel<T, true, Q, 4>), bps, tpb, 0, s, A, zc, ntx, nty); \
            else hipLaunchKernelGGL((adjoint_transpose_lds_kernel<T, false, Q, 4>), bps, tpb, 0, s, A, zc, ntx, nty);    \
        }                                                                                          \
        else if (upwd) hipLaunchKernelGGL((adjoint_transpose_z_kernel<T, true, Q>), bpz, tpb, 0, s, A, zc); \
        else hipLaunchKernelGGL((adjoint_transpose_z_kernel<T, false, Q>), bpz, tpb, 0, s, A, zc);    \
    } while (0)
    AC_ADJ_Q(0);
    AC_ADJ_Q(1);
    AC_ADJ_Q(2);
    AC_ADJ_Q(3);
    AC_ADJ_Q(4);
    AC_ADJ_Q(5);
    AC_ADJ_Q(6);
    // without the entropy equation ss is not evolved (f'_7 = f_7): its lambda_f passes through
    // and the local stage has set its lambda_w to alpha * 0
    if (entropy)
        AC_ADJ_Q(7);
#undef AC_ADJ_Q
#undef AC_ADJ
}

int adjoint_num_fields() { return NG; }

template void launch_tangent<AcRealT>(const AdjArgs<AcRealT>&, bool, bool, hipStream_t);
template void launch_adjoint<AcRealT>(const AdjArgs<AcRealT>&, bool, bool, double*, hipStream_t);

}  // namespace acamd

// ------------------------------------------------------------------ C entry points
#include "../../include/astaroth.h"

namespace acamd {
// defined in device.hip
void device_solve_args(Device d, SolveArgs<AcRealT>& a, int step, AcRealT dt);
hipStream_t device_stream(Device d, Stream s);
int device_id(Device d);
}  // namespace acamd

using namespace acamd;

extern "C" size_t acAdjointWorkspaceBytes(const Device d)
{
    const AcMeshInfo info = acDeviceGetLocalConfig(d);
    const int3 m = info.int3_params[AC_mlocal];
    const int3 n = info.int3_params[AC_nlocal];
    const size_t count = (size_t)m.x * m.y * m.z;
    const size_t g = count * NG * sizeof(AcReal);
    const size_t partials =
        ((size_t)adjoint_num_blocks(n.x, n.y, n.z) + PARAM_RANGES) * 2 * sizeof(double);
    return ((g + 255) / 256) * 256 + partials;
}

extern "C" AcResult acDeviceIntegrateSubstepTangent(const Device d, const Stream stream,
                                                    const int step_number, const AcReal dt,
                                                    const AcReal* df_in, AcReal* df_out,
                                                    AcReal* dw, const double* dparams)
{
    if (step_number < 0 || step_number > 2) {
        fprintf(stderr, "astaroth_amd: acDeviceIntegrateSubstepTangent: step_number must be 0..2\n");
        return AC_FAILURE;
    }
    if (hipSetDevice(device_id(d)) != hipSuccess)
        return AC_FAILURE;
    const AcMeshInfo info = acDeviceGetLocalConfig(d);
    const int3 m = info.int3_params[AC_mlocal];
    AdjArgs<AcReal> A;
    memset(&A, 0, sizeof(A));
    device_solve_args(d, A.s, step_number, dt);
    A.count = (long long)m.x * m.y * m.z;
    A.df_in = df_in;
    A.df_out = df_out;
    A.dw = dw;
    A.dnu = dparams ? (AcReal)dparams[0] : (AcReal)0;
    A.deta = dparams ? (AcReal)dparams[1] : (AcReal)0;
    hipStream_t s = device_stream(d, stream);
    AcReal* ptrs[NF];
    for (int q = 0; q < NF; ++q)
        ptrs[q] = const_cast<AcReal*>(df_in) + (long long)q * A.count;
    launch_periodic_fill<AcReal>(ptrs, NF, A.s.nx, A.s.ny, A.s.nz, s);
    launch_tangent<AcReal>(A, info.bool_params[AC_lupwd] != 0, info.bool_params[AC_lentropy] != 0, s);
    return hipGetLastError() == hipSuccess ? AC_SUCCESS : AC_FAILURE;
}

extern "C" AcResult acDeviceIntegrateSubstepAdjoint(const Device d, const Stream stream,
                                                    const int step_number, const AcReal dt,
                                                    AcReal* lam_f, AcReal* lam_w, void* workspace,
                                                    double* param_grad)
{
    if (step_number < 0 || step_number > 2) {
        fprintf(stderr, "astaroth_amd: acDeviceIntegrateSubstepAdjoint: step_number must be 0..2\n");
        return AC_FAILURE;
    }
    if (hipSetDevice(device_id(d)) != hipSuccess)
        return AC_FAILURE;
    const AcMeshInfo info = acDeviceGetLocalConfig(d);
    const int3 m = info.int3_params[AC_mlocal];
    AdjArgs<AcReal> A;
    memset(&A, 0, sizeof(A));
    device_solve_args(d, A.s, step_number, dt);
    A.count = (long long)m.x * m.y * m.z;
    A.lf = lam_f;
    A.lw = lam_w;
    A.g = (AcReal*)workspace;
    const size_t gbytes = ((size_t)A.count * NG * sizeof(AcReal) + 255) / 256 * 256;
    A.partials = (double*)((char*)workspace + gbytes);
    hipStream_t s = device_stream(d, stream);
    launch_adjoint<AcReal>(A, info.bool_params[AC_lupwd] != 0, info.bool_params[AC_lentropy] != 0,
                           param_grad, s);
    return hipGetLastError() == hipSuccess ? AC_SUCCESS : AC_FAILURE;
}
