// solve_tiled.hip -- LDS-tiled, register-rolled (2.5D) fused RK3 substep for gfx950.
//
// Same semantics as solve_direct.hip (ref acc-runtime/samples/mhd_modular/mhdsolver.ac:586-636,
// one fused pass), organised for HBM reuse:
//
//  * A workgroup owns a TX x TY column block of the interior and a chunk of ZC planes. It
//    marches upwards in z; each step stages one xy plane of all eight fields, with its radius-3
//    halo, in LDS ((TX+6) x (TY+6) values per field), loaded with wave64-coalesced row reads and
//    software-prefetched one plane ahead through registers.
//  * Every thread owns one (x, y) column. From the plane in LDS it takes its in-plane stencils
//    (x and y first, second, upwind and xy-mixed derivatives). Everything that needs other
//    planes stays in registers:
//      - a 7-deep queue of its own column's values (z first/second/upwind derivatives);
//      - "chains" for the xz / yz mixed derivatives, written as
//            d_xz f(z) = sum_i c_i [ Dx_i f(z+i) - Dx_i f(z-i) ],  Dx_i f = f(x+i) - f(x-i),
//        where each arriving plane p pushes +-c_i Dx_i f(p) into the accumulators of planes
//        p -+ i. Only the combinations grad(div u) and grad(div A) are accumulated, with their
//        in-plane parts (d_xx u_x + d_xy u_y, ...) added when the plane itself arrives;
//      - a 3-plane delay line of the in-plane derivatives, because the output plane z lags the
//        newest plane p = z + 3.
//  * No MFMA: a stencil is not a dense contraction. The kernel is bound by HBM traffic and VALU.
//  * Workgroups are numbered so that consecutive tiles (which share halo rows) land on the same
//    XCD and its L2: hardware dispatches workgroup b to XCD b % 8.
#include "kernels.h"
#include "mhd_adjoint.h"

namespace acamd {

// In-plane quantities of one plane that the output needs three planes later. The lnrho and u
// ones ride a register delay line; the ten of A and ss (ax_dy, ay_dx, az_dx, az_dy, a_lap[3],
// s_dx, s_dy, s_lap, in that order) a per-thread three-slot ring in LDS, which keeps the
// register footprint within the budget of two waves per SIMD.
template <typename T>
struct PlaneVals {
    T lr_dx, lr_dy, lr_lap, lr_up;       // lnrho: d_x, d_y, d_xx + d_yy, |ux| up_x + |uy| up_y
    T u_dx[3], u_dy[3], u_lap[3], u_up[3];
    T lr_upx, lr_upy, u_upx[3], u_upy[3];  // upwind terms per direction (adjoint body only)
};
constexpr int NDL = 10;  // delayed values kept in LDS

// Chains: 0..2 grad(div u) x, y, z ; 3..5 grad(div A) x, y, z.
constexpr int NCH = 6;

template <typename T, int LX>
struct Lds {
    const T* p;  // this thread's centre in the plane of one field
    __device__ __forceinline__ T at(int dx, int dy) const { return p[dy * LX + dx]; }
};

// The plane loop shared by the forward solve (ADJ = false: right-hand side and RK3 update of the
// output cell) and the adjoint's local stage (ADJ = true: derivative adjoints of the output cell
// from the same derivative quantities, with mu = lambda_w + beta lambda_f; see adjoint.hip).
template <typename T, bool UPWD, bool ENTROPY, int KB, int TX, int TY, bool ADJ, bool WRAP, int STEP>
__device__ __forceinline__ void delay_line_body(const SolveArgs<T>& a, const AdjLocal<T>& x,
                                                int zchunk, int ntx, int nty, int nchunks)
{
    constexpr int LX = TX + 2 * R;
    constexpr int LY = TY + 2 * R;
    constexpr int LP = LX * LY;                  // one field-plane with halo
    constexpr int NT = TX * TY;
    constexpr int NW = NT / 64;                  // waves per workgroup
    constexpr int WPE = (int)sizeof(T) / 4;      // dwords per element
    constexpr int LPW = LP * WPE;                // dwords per field-plane
    constexpr int NCHUNK = (LPW + 63) / 64;      // wave-wide DMA chunks (one dword per lane)
    constexpr int NR = (NCHUNK + NW - 1) / NW;   // chunks per wave per field-plane
    // Each field-plane occupies NR * NW whole chunks in LDS, so every wave issues exactly NR
    // DMA instructions per field with all lanes active (no exec-mask branches): lanes past the
    // plane's last dword read a valid address and land in the padding.
    constexpr int LPS = NR * NW * 64 / WPE;      // LDS stride of one field-plane (elements)
    // The precise FP32 mode (32-bit compensation): second, upwind and mixed derivatives from
    // differences to the centre value (exact for neighbouring values) instead of sums of
    // neighbours, whose rounding in FP32 is of order eps |f| / dx^2 against a result of order
    // k^2 |f|; about 10x smaller errors in u, A and lnrho for about 18% more VALU work
    constexpr bool CANCEL_FREE = sizeof(T) == 4 && !ADJ && KB == 32;
    const int step = STEP >= 0 ? STEP : a.step;  // the RK3 substep (forward body)
    // LATE: the output of plane p - 3 (with its stores) runs after the end-of-step barrier, and
    // the DMA runs two planes ahead, issued right after that barrier into the buffer just
    // released. The barrier's wait then finds the previous step's stores and the next plane's
    // DMA long done, instead of waiting for the stores just issued (vmcnt cannot tell loads from
    // stores on this target, so the wait before the barrier covers both). Used in FP32 by the
    // adjoint's local stage (47 derivative adjoints, lambda_w and lambda_f stored per cell) and
    // by the forward kernels up to 16-bit compensation (plain FP32 0.598 -> 0.584 ms per launch
    // at 256^3, 16-bit 0.68 -> 0.66 ms);
    // the 32-bit compensated and FP64 kernels would spill the longer-lived registers.
    constexpr bool LATE = sizeof(T) == 4 && (ADJ || KB < 32);
    // Skip the in-plane stage (3) on the planes that are not output. Measured: FP32 plain, 8- and
    // 16-bit compensated kernels 0-5% faster; the 32-bit compensated and FP64 kernels 1.5-2%
    // slower (different register allocation), so they keep the stage on every plane.
    constexpr bool SKIP_EDGE = sizeof(T) == 4 && KB < 32;
    static_assert(NT % 64 == 0, "workgroup must be whole waves");
    static_assert((NR * NW * 64) % WPE == 0, "field-plane stride");
    __shared__ T lds[2][NF * LPS];
    __shared__ T dl[3][NDL][NT];  // delayed A / ss values, [slot][value][thread]

    // ---- XCD-aware tile numbering: workgroup b runs on XCD b % 8; give each XCD a contiguous
    // range of tiles (x fastest, then y, then z-chunk) so halo rows are shared within one L2.
    const int nblocks = ntx * nty * nchunks;
    const int bid = blockIdx.x;
    int logical = bid;
    if ((nblocks & 7) == 0)
        logical = (bid & 7) * (nblocks >> 3) + (bid >> 3);
    int tile_x = logical % ntx;
    // With two workgroups per CU the dispatcher places workgroups b and b + nblocks / 2 on one
    // CU (observed with per-workgroup clocks at 256^3). The tiles at the x edges, which read
    // periodic images across the boundary, took 2.5% (plain) to 5.7% (16-bit compensated)
    // longer, so CUs holding two of them set the launch time. In the second half tile_x ^ 1
    // pairs each edge tile with an inner one of the same rows: a bijection when each XCD's
    // half-range is whole rows of an even number of tiles (checked here; otherwise unchanged).
    if (sizeof(T) == 4 && !ADJ && (ntx & 1) == 0 && (nblocks & 15) == 0 &&
        (nblocks >> 4) % ntx == 0 && 2 * bid >= nblocks)
        tile_x ^= 1;
    const int rest = logical / ntx;
    const int tile_y = rest % nty;
    const int chunk = rest / nty;

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = tx + TX * ty;
    // The wave index is uniform across the wave. FP32: kept in a scalar register, so the LDS
    // destinations of the DMA (M0) are scalar arithmetic (about 70 fewer vector instructions per
    // plane step). FP64 keeps the per-DMA vector form: its kernels run at 512 VGPRs with spills,
    // and with the scalar form the compiled periodic-image kernels computed wrong results (64 x 2
    // tiles in one build, 64 x 4 in another) although the source indexes the same LDS and global
    // elements either way, and their assembly reads no undefined register, spill slot or lane.
    const int wave = sizeof(T) == 4 ? __builtin_amdgcn_readfirstlane(tid >> 6) : tid >> 6;
    const int lane = tid & 63;
    // Two workgroups share each CU in FP32, one wave of each per SIMD, and the SIMD's arbiter
    // favours the wave in slot 0 when both can issue: measured per workgroup, the slot-0 one
    // finished about 7% earlier and the other ran its last planes alone, at half occupancy.
    // The slot-1 wave takes the higher issue priority on every other plane step, so the two
    // progress at about the same rate (256^3: 0.5846 -> 0.5652-0.5686 ms per launch). Slot =
    // HW_ID bits 3:0, read once (wave-uniform).
    constexpr bool PRIO = sizeof(T) == 4 && !ADJ;
    const unsigned wslot = PRIO ? (__builtin_amdgcn_s_getreg((3 << 11) | 4) & 1u) : 0u;
    const int x0 = tile_x * TX, y0 = tile_y * TY, z0 = chunk * zchunk;
    const long long sy = a.mx, sz = a.sz;

    // ---- DMA slots of this lane: chunk j = wave + r * NW covers the dwords [64 j, 64 j + 64) of
    // a field-plane; lane l moves dword d = 64 j + l, i.e. word d % WPE of element d / WPE.
    unsigned g_off[NR];  // byte offset of that dword within the array plane
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int j = wave + r * NW;
        int d = j * 64 + lane;
        if (d >= LPW)
            d = 0;  // padding lane: re-read the first dword (in bounds, value unused)
        const int e = d / WPE, wd = d - e * WPE;
        const int ly = e / LX, lx = e - ly * LX;
        int hx = x0 + lx, hy = y0 + ly;  // array coordinates of the element
        if (WRAP) {
            // halo elements from their periodic images in the interior
            hx += hx < R ? a.nx : (hx >= a.nx + R ? -a.nx : 0);
            hy += hy < R ? a.ny : (hy >= a.ny + R ? -a.ny : 0);
        }
        g_off[r] = (unsigned)((hx + (int)sy * hy) * (int)sizeof(T) + 4 * wd);
    }
    // Field f of a role lives at slab + f * fstride; the per-field bases are recomputed from a
    // stride made opaque once per plane step (scalar work) rather than held across the loop.
    const char* const in0 = (const char*)a.in[0];
    char* const out0 = (char*)a.out[0];
    char* const w0 = (char*)a.w[0];
    char* const comp0 = (char*)a.comp[0];
    // Per-field bases of the RK arrays. With 16- and 32-bit compensation they are pinned to
    // scalar registers (sbase), which keeps those kernels free of vector spills (about 3% and
    // 20% faster at 256^3); the plain and 8-bit kernels are faster with the compiler's own choice.
    constexpr bool SBASE = KB >= 16;
    auto field_base = [](char* p) {
        if constexpr (SBASE)
            return sbase(p);
        else
            return (T*)p;
    };
    auto ld_field = [](char* p, unsigned boff) -> T {
        if constexpr (SBASE)
            return ld_g<T>(sbase(p), boff);
        else
            return ld_b((const T*)p, boff);
    };
    // Issue the DMA of interior plane p (array z index p + R) into buffer b.
    auto dma_plane = [&](int p, int b, long long fsb) {
        if (WRAP)
            p += p < 0 ? a.nz : (p >= a.nz ? -a.nz : 0);
        const long long base = sz * (long long)(p + R) * (long long)sizeof(T);
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const char* src = in0 + (base + f * fsb);
            char* dst = (char*)&lds[b][f * LPS];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const int j = wave + r * NW;
                __builtin_amdgcn_global_load_lds((const void*)(src + opaque(g_off[r])),
                                                 (lds_ptr_t)(dst + 256 * j), 4, 0, 0);
            }
        }
    };

    const Coeffs<T>& C = a.C;
    T q[NF][7];
    T ch[NCH][7];
    // The delay line starts zeroed like the queues: every register the loop carries is then
    // written before it is read on every path (with it left uninitialized, the compiled FP64
    // kernels copied never-written register pairs through their AGPR spill space each plane
    // step; see the ISA check in tests/test_isa_invariants.py).
    PlaneVals<T> dly[4] = {};
    int dslot = 0;  // ring slot of this step's plane p (and of the output plane p - 3)
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int i = 0; i < 7; ++i)
            q[f][i] = T(0);
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
        for (int i = 0; i < 7; ++i)
            ch[c][i] = T(0);

    const T dt = solve_dt(a);
    double u2max = 0.0;  // max |u_new|^2 (the courant timestep mode)
    double pnu = 0.0, peta = 0.0;  // adjoint body: dt <mu, dR/dnu>, dt <mu, dR/deta>
    const int gx = x0 + tx, gy = y0 + ty;
    const int ctr = (ty + R) * LX + (tx + R);
    const unsigned ioff = (unsigned)((gx + R) + (int)sy * (gy + R));  // in-plane offset

    dma_plane(z0 - R, 0, a.fstride * (long long)sizeof(T));
    if (LATE)
        dma_plane(z0 - R + 1, 1, a.fstride * (long long)sizeof(T));
    plane_barrier();
    const int niter = zchunk + 2 * R;
    for (int it = 0; it < niter; ++it) {
        const int p = z0 - R + it;
        const int cur = it & 1;
        if constexpr (PRIO) {
            if (wslot & (unsigned)it)
                __builtin_amdgcn_s_setprio(1);
            else
                __builtin_amdgcn_s_setprio(0);
        }
        const unsigned iob = opaque(ioff * (unsigned)sizeof(T));  // byte offset in a plane
        const long long fsb = opaque_s64(a.fstride * (long long)sizeof(T));  // field stride, bytes
        if (!LATE && it + 1 < niter)
            dma_plane(p + 1, cur ^ 1, fsb);
        // w of this iteration's output cell, loaded early so its latency hides behind the
        // stencil work (it is needed only by the RK update at the end)
        T wpre[NF];
        CompRec<KB> crec{};  // the output cell's compensation record (KB > 0)
        T lfpre[NF];  // adjoint body: lambda_f and lambda_w of the output cell
        const int zo = p - R;
        const bool has_out = it >= 2 * R;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            wpre[f] = T(0);
            lfpre[f] = T(0);
        }
        if (ADJ && has_out) {
            const long long poff = sz * (long long)(zo + R);
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const long long fo = f * fsb + poff * (long long)sizeof(T);
                lfpre[f] = ld_b((const T*)((const char*)x.lf + fo), iob);
                wpre[f] = ld_b((const T*)((const char*)x.lw + fo), iob);
            }
        }
        if (!ADJ && has_out) {
            const long long poff = sz * (long long)(zo + R);
            if (step > 0) {
#pragma unroll
                for (int f = 0; f < NF; ++f)
                    wpre[f] = ld_field(w0 + (f * fsb + poff * (long long)sizeof(T)), iob);
            }
            if (KB > 0)
                crec = comp_rec_load<KB>(comp0 + poff * (long long)(NF * KB / 8),
                                         iob / (unsigned)sizeof(T) * (unsigned)(NF * KB / 8));
        }

        Lds<T, LX> L[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            L[f].p = &lds[cur][f * LPS + ctr];
            q[f][6] = L[f].at(0, 0);
        }
        PlaneVals<T>& nv = dly[3];
        T az_dx_p = T(0), az_dy_p = T(0);  // of plane p, stored to the LDS ring below

        // ---- (1) z-chains: plane p adds +c_i D(p) to plane p - i and -c_i D(p) to plane p + i
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const int fx = 1 + 3 * v, fy = 2 + 3 * v, fz = 3 + 3 * v;
            T dxz = T(0), dyz = T(0), dxx_ = T(0), dyy_ = T(0);
#pragma unroll
            for (int i = 1; i <= 3; ++i) {
                const T Dxz = L[fz].at(i, 0) - L[fz].at(-i, 0);
                const T Dyz = L[fz].at(0, i) - L[fz].at(0, -i);
                const T Dxx = L[fx].at(i, 0) - L[fx].at(-i, 0);
                const T Dyy = L[fy].at(0, i) - L[fy].at(0, -i);
                const T kx = C.mx[1][i], ky = C.mx[2][i];
                // d_xz v_z -> (grad div v)_x, d_yz v_z -> (grad div v)_y,
                // d_xz v_x + d_yz v_y -> (grad div v)_z
                const T ez = kx * Dxx + ky * Dyy;
                if constexpr (sizeof(T) == 4) {
                    // one FMA per chain entry (FP64 keeps the code its register allocation
                    // was measured with: this form made it 6% slower)
                    ch[3 * v + 0][3 - i] = ac_fma(kx, Dxz, ch[3 * v + 0][3 - i]);
                    ch[3 * v + 0][3 + i] = ac_fma(-kx, Dxz, ch[3 * v + 0][3 + i]);
                    ch[3 * v + 1][3 - i] = ac_fma(ky, Dyz, ch[3 * v + 1][3 - i]);
                    ch[3 * v + 1][3 + i] = ac_fma(-ky, Dyz, ch[3 * v + 1][3 + i]);
                }
                else {
                    const T ex = kx * Dxz, ey = ky * Dyz;
                    ch[3 * v + 0][3 - i] += ex;
                    ch[3 * v + 0][3 + i] -= ex;
                    ch[3 * v + 1][3 - i] += ey;
                    ch[3 * v + 1][3 + i] -= ey;
                }
                ch[3 * v + 2][3 - i] += ez;
                ch[3 * v + 2][3 + i] -= ez;
                dxz += C.d1[0][i] * Dxz;
                dyz += C.d1[1][i] * Dyz;
                dxx_ += C.d1[0][i] * Dxx;
                dyy_ += C.d1[1][i] * Dyy;
            }
            if (v == 0) {
                nv.u_dx[2] = dxz;
                nv.u_dy[2] = dyz;
                nv.u_dx[0] = dxx_;
                nv.u_dy[1] = dyy_;
            }
            else {
                az_dx_p = dxz;
                az_dy_p = dyz;
                (void)dxx_;
                (void)dyy_;
            }
        }

        // ---- (2) output plane z = p - 3: here, or after the end-of-step barrier when LATE
        const int z = p - R;
        T dv[NDL];  // the A / ss values of plane z, written three steps ago
        auto read_dv = [&]() __attribute__((always_inline)) {
#pragma unroll
            for (int k = 0; k < NDL; ++k)
                dv[k] = dl[dslot][k][tid];
        };
        auto output_phase = [&]() __attribute__((always_inline)) {
            const PlaneVals<T>& o = dly[0];
            T dz[NF], dzz[NF], upz[4];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                T D[4], S[4];
#pragma unroll
                for (int i = 1; i <= 3; ++i) {
                    D[i] = q[f][3 + i] - q[f][3 - i];
                    S[i] = CANCEL_FREE ? (q[f][3 + i] - q[f][3]) + (q[f][3 - i] - q[f][3])
                                       : q[f][3 + i] + q[f][3 - i];
                }
                dz[f] = C.d1[2][1] * D[1] + C.d1[2][2] * D[2] + C.d1[2][3] * D[3];
                if constexpr (CANCEL_FREE) {
                    dzz[f] = C.d2[2][1] * S[1] + C.d2[2][2] * S[2] + C.d2[2][3] * S[3];
                    if (UPWD && f < 4)
                        upz[f] = C.up[2][1] * S[1] + C.up[2][2] * S[2] + C.up[2][3] * S[3];
                }
                else {
                    dzz[f] = C.d2[2][0] * q[f][3] + C.d2[2][1] * S[1] + C.d2[2][2] * S[2] + C.d2[2][3] * S[3];
                    if (UPWD && f < 4)
                        upz[f] = C.up[2][0] * q[f][3] + C.up[2][1] * S[1] + C.up[2][2] * S[2] + C.up[2][3] * S[3];
                }
            }
            CellDerivs<T> d;
            d.lr = q[0][3];
            d.glr[0] = o.lr_dx;
            d.glr[1] = o.lr_dy;
            d.glr[2] = dz[0];
            d.laplr = o.lr_lap + dzz[0];
            const T auz = ac_abs(q[3][3]);
            d.upl = (UPWD && !ADJ) ? o.lr_up + auz * upz[0] : T(0);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                d.u[i] = q[1 + i][3];
                d.gu[i][0] = o.u_dx[i];
                d.gu[i][1] = o.u_dy[i];
                d.gu[i][2] = dz[1 + i];
                d.lapu[i] = o.u_lap[i] + dzz[1 + i];
                d.upu[i] = (UPWD && !ADJ) ? o.u_up[i] + auz * upz[1 + i] : T(0);
                d.lapa[i] = dv[4 + i] + dzz[4 + i];
            }
            d.gdu[0] = ch[0][0];
            d.gdu[1] = ch[1][0];
            d.gdu[2] = ch[2][0] + dzz[3];
            d.gda[0] = ch[3][0];
            d.gda[1] = ch[4][0];
            d.gda[2] = ch[5][0] + dzz[6];
            d.ga[0][0] = T(0);
            d.ga[0][1] = dv[0];
            d.ga[0][2] = dz[4];
            d.ga[1][0] = dv[1];
            d.ga[1][1] = T(0);
            d.ga[1][2] = dz[5];
            d.ga[2][0] = dv[2];
            d.ga[2][1] = dv[3];
            d.ga[2][2] = T(0);
            d.s = q[7][3];
            d.gs[0] = dv[7];
            d.gs[1] = dv[8];
            d.gs[2] = dz[7];
            d.laps = dv[9] + dzz[7];

            if constexpr (ADJ) {
                UpwindDirs<T> up;
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    up.ul[j] = T(0);
#pragma unroll
                    for (int i = 0; i < 3; ++i)
                        up.uu[i][j] = T(0);
                }
                if (UPWD) {
                    up.ul[0] = o.lr_upx;
                    up.ul[1] = o.lr_upy;
                    up.ul[2] = upz[0];
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        up.uu[i][0] = o.u_upx[i];
                        up.uu[i][1] = o.u_upy[i];
                        up.uu[i][2] = upz[1 + i];
                    }
                    combine_upwind(d, up);
                }
                T mu[NF];
#pragma unroll
                for (int f = 0; f < NF; ++f)
                    mu[f] = wpre[f] + a.beta * lfpre[f];
                if (!ENTROPY)
                    mu[7] = T(0);
                CellDerivs<T> ad;
                UpwindDirs<T> aup;
                T anu, aeta;
                mhd_rhs_vjp<T, UPWD, ENTROPY>(d, up, mu, a.P, ad, aup, anu, aeta);
                pnu += (double)dt * (double)anu;
                peta += (double)dt * (double)aeta;
                const long long poffb = sz * (long long)(z + R) * (long long)sizeof(T);
                put_derivative_adjoints<UPWD>(ad, aup, [&](int k, T v) {
                    st_b((T*)((char*)x.g + (k * fsb + poffb)), iob, v);
                });
                // the pointwise parts of the substep's adjoint: lambda_w <- alpha mu, and
                // lambda_f gains dt times the adjoint of the cell's own value (the transposed
                // stencils add the rest)
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    st_b((T*)((char*)x.lw + (f * fsb + poffb)), iob, a.alpha * mu[f]);
                    if (f < 4 || (f == 7 && ENTROPY))
                        st_b((T*)((char*)x.lf + (f * fsb + poffb)), iob,
                             lfpre[f] + dt * value_adjoint(ad, f));
                }
            }
            else {
            T rhs[NF];
            mhd_rhs<T, UPWD, ENTROPY, CANCEL_FREE>(d, a.P, rhs);
            const long long poff = sz * (long long)(z + R);
            T fnew[NF];
            CompRec<KB> cout = crec;  // ss keeps its record entry without entropy
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const long long fo = f * fsb + poff * (long long)sizeof(T);  // bytes
                auto op = field_base(out0 + fo);
                if (f == 7 && !ENTROPY) {
                    fnew[f] = q[f][3];
                    st_b(op, iob, q[f][3]);
                    continue;
                }
                auto wp = field_base(w0 + fo);
                int cenc = 0;
                fnew[f] = rk_update_pre<T, KB>(wp, op, iob, step, a.alpha, a.beta, q[f][3], rhs[f],
                                               dt, wpre[f], comp_rec_get<KB>(crec, f), &cenc);
                if (KB > 0)
                    comp_rec_set<KB>(cout, f, cenc);
            }
            if (KB > 0)
                comp_rec_store<KB>(comp0 + poff * (long long)(NF * KB / 8),
                                   iob / (unsigned)sizeof(T) * (unsigned)(NF * KB / 8), cout);
            if (step == 2) {
                const double u2 = (double)fnew[1] * (double)fnew[1] + (double)fnew[2] * (double)fnew[2] +
                                  (double)fnew[3] * (double)fnew[3];
                u2max = fmax(u2max, u2);
            }
            }
        };
        if constexpr (!LATE) {
            if (it >= 2 * R) {
                read_dv();
                output_phase();
            }
        }
        else if (it >= 2 * R) {
            read_dv();  // before (3) overwrites the ring slot with plane p
        }


        // ---- (3) remaining in-plane derivatives of plane p, needed only where plane p is itself
        // output (planes z0 .. z0 + zchunk - 1; the first and last three of a chunk only feed the
        // z-stencils and chains of their neighbours)
        if (!SKIP_EDGE || (it >= R && it < zchunk + R)) {
            const T aux = ac_abs(q[1][6]), auy = ac_abs(q[2][6]);
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const T c0 = q[f][6];
                T Dx[4], Dy[4], Sx[4], Sy[4];
#pragma unroll
                for (int i = 1; i <= 3; ++i) {
                    const T xp = L[f].at(i, 0), xm = L[f].at(-i, 0);
                    const T yp = L[f].at(0, i), ym = L[f].at(0, -i);
                    Dx[i] = xp - xm;
                    Dy[i] = yp - ym;
                    if constexpr (CANCEL_FREE) {
                        Sx[i] = (xp - c0) + (xm - c0);
                        Sy[i] = (yp - c0) + (ym - c0);
                    }
                    else {
                        Sx[i] = xp + xm;
                        Sy[i] = yp + ym;
                    }
                }
                T upxy = T(0), upx = T(0), upy = T(0);
                T dxx, dyy;
                if constexpr (CANCEL_FREE) {
                    // S holds f(+i) + f(-i) - 2 f(0): the centre weight (minus twice the sum of the
                    // others) drops out
                    dxx = C.d2[0][1] * Sx[1] + C.d2[0][2] * Sx[2] + C.d2[0][3] * Sx[3];
                    dyy = C.d2[1][1] * Sy[1] + C.d2[1][2] * Sy[2] + C.d2[1][3] * Sy[3];
                    if (UPWD && f < 4) {
                        upx = C.up[0][1] * Sx[1] + C.up[0][2] * Sx[2] + C.up[0][3] * Sx[3];
                        upy = C.up[1][1] * Sy[1] + C.up[1][2] * Sy[2] + C.up[1][3] * Sy[3];
                        upxy = aux * upx + auy * upy;
                    }
                }
                else {
                dxx = C.d2[0][0] * c0 + C.d2[0][1] * Sx[1] + C.d2[0][2] * Sx[2] + C.d2[0][3] * Sx[3];
                dyy = C.d2[1][0] * c0 + C.d2[1][1] * Sy[1] + C.d2[1][2] * Sy[2] + C.d2[1][3] * Sy[3];
                if (UPWD && f < 4) {
                    upx = C.up[0][0] * c0 + C.up[0][1] * Sx[1] + C.up[0][2] * Sx[2] + C.up[0][3] * Sx[3];
                    upy = C.up[1][0] * c0 + C.up[1][1] * Sy[1] + C.up[1][2] * Sy[2] + C.up[1][3] * Sy[3];
                    if (!ADJ)
                        upxy = aux * upx + auy * upy;
                }
                }
                const bool need_dx = (f == 0 || f == 2 || f == 5 || f == 7);
                const bool need_dy = (f == 0 || f == 1 || f == 4 || f == 7);
                T dx = T(0), dy = T(0);
                if (need_dx)
                    dx = C.d1[0][1] * Dx[1] + C.d1[0][2] * Dx[2] + C.d1[0][3] * Dx[3];
                if (need_dy)
                    dy = C.d1[1][1] * Dy[1] + C.d1[1][2] * Dy[2] + C.d1[1][3] * Dy[3];
                if (f == 0) {
                    nv.lr_dx = dx;
                    nv.lr_dy = dy;
                    nv.lr_lap = dxx + dyy;
                    nv.lr_up = upxy;
                    if (ADJ) {
                        nv.lr_upx = upx;
                        nv.lr_upy = upy;
                    }
                }
                else if (f < 4) {
                    if (f == 1) nv.u_dy[0] = dy;
                    if (f == 2) nv.u_dx[1] = dx;
                    nv.u_lap[f - 1] = dxx + dyy;
                    nv.u_up[f - 1] = upxy;
                    if (ADJ) {
                        nv.u_upx[f - 1] = upx;
                        nv.u_upy[f - 1] = upy;
                    }
                }
                else if (f < 7) {
                    if (f == 4) dl[dslot][0][tid] = dy;
                    if (f == 5) dl[dslot][1][tid] = dx;
                    dl[dslot][4 + (f - 4)][tid] = dxx + dyy;
                }
                else {
                    dl[dslot][7][tid] = dx;
                    dl[dslot][8][tid] = dy;
                    dl[dslot][9][tid] = dxx + dyy;
                }
                // in-plane parts of grad(div v): x gets d_xx v_x, y gets d_yy v_y
                if (f == 1) ch[0][3] += dxx;
                if (f == 2) ch[1][3] += dyy;
                if (f == 4) ch[3][3] += dxx;
                if (f == 5) ch[4][3] += dyy;
            }
            // xy-mixed derivatives of u_x, u_y, A_x, A_y
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const int fx = 1 + 3 * v, fy = 2 + 3 * v;
                T mxx = T(0), myy = T(0);
#pragma unroll
                for (int i = 1; i <= 3; ++i) {
                    if constexpr (CANCEL_FREE) {
                        mxx += C.mx[0][i] * ((L[fx].at(i, i) - L[fx].at(i, -i)) - (L[fx].at(-i, i) - L[fx].at(-i, -i)));
                        myy += C.mx[0][i] * ((L[fy].at(i, i) - L[fy].at(i, -i)) - (L[fy].at(-i, i) - L[fy].at(-i, -i)));
                    }
                    else {
                        mxx += C.mx[0][i] * ((L[fx].at(i, i) + L[fx].at(-i, -i)) - (L[fx].at(i, -i) + L[fx].at(-i, i)));
                        myy += C.mx[0][i] * ((L[fy].at(i, i) + L[fy].at(-i, -i)) - (L[fy].at(i, -i) + L[fy].at(-i, i)));
                    }
                }
                ch[3 * v + 0][3] += myy;  // d_xy v_y into (grad div v)_x
                ch[3 * v + 1][3] += mxx;  // d_xy v_x into (grad div v)_y
            }
        }

        if constexpr (LATE) {
            // every wave is done with buffer cur (plane p) and plane p + 1 has landed in the
            // other one; the DMA of plane p + 2 into cur then overlaps the output work, and the
            // output's stores have the next step's stencil work to complete before its barrier
            plane_barrier();
            if (it + 2 < niter)
                dma_plane(p + 2, cur, fsb);
            if (it >= 2 * R)
                output_phase();
        }

        // ---- (4) roll the register state by one plane
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int i = 0; i < 6; ++i)
                q[f][i] = q[f][i + 1];
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
#pragma unroll
            for (int i = 0; i < 6; ++i)
                ch[c][i] = ch[c][i + 1];
            ch[c][6] = T(0);
        }
        dl[dslot][2][tid] = az_dx_p;
        dl[dslot][3][tid] = az_dy_p;
        dly[0] = dly[1];
        dly[1] = dly[2];
        dly[2] = dly[3];
        dslot = dslot == 2 ? 0 : dslot + 1;

        if constexpr (!LATE) {
            // plane p + 1 has landed in the other buffer and every wave is done with this one
            plane_barrier();
        }
    }
    if constexpr (ADJ) {
        // block sums of the parameter sensitivities (FP64, deterministic)
        __shared__ double psh[2][NW];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            pnu += __shfl_xor(pnu, off, 64);
            peta += __shfl_xor(peta, off, 64);
        }
        if (lane == 0) {
            psh[0][wave] = pnu;
            psh[1][wave] = peta;
        }
        __syncthreads();
        if (tid == 0) {
            double s0 = 0.0, s1 = 0.0;
            for (int w = 0; w < NW; ++w) {
                s0 += psh[0][w];
                s1 += psh[1][w];
            }
            x.partials[2 * (long long)blockIdx.x] = s0;
            x.partials[2 * (long long)blockIdx.x + 1] = s1;
        }
    }
    else if (step == 2) {
        if (a.umax2_next != nullptr)
            wave_max_to_global(u2max, a.umax2_next);
        if (a.time_acc != nullptr && blockIdx.x == 0 && tid == 0)
            *a.time_acc += (double)dt;
    }
}

// WRAP: read the halo from the periodic images inside the interior (a.wrap), a separate
// instantiation so that the plain kernels keep their own code generation. STEP: the RK3 substep
// (0..2) fixed at compile time, or -1 to read a.step.
template <typename T, bool UPWD, bool ENTROPY, int KB, int TX, int TY, bool WRAP, int STEP>
__global__ void __launch_bounds__(TX* TY, (sizeof(T) == 4 && TX * TY == 256 ? 2 : 1))
    solve_tiled_kernel(SolveArgs<T> a, int zchunk, int ntx, int nty, int nchunks)
{
    delay_line_body<T, UPWD, ENTROPY, KB, TX, TY, false, WRAP, STEP>(a, AdjLocal<T>{}, zchunk, ntx,
                                                                     nty, nchunks);
}

// The adjoint's local stage on 64 x 4 tiles, one wave per SIMD (the VJP needs the registers).
template <typename T, bool UPWD, bool ENTROPY>
__global__ void __launch_bounds__(256, 1)
    adjoint_local_tiled_kernel(SolveArgs<T> a, AdjLocal<T> x, int zchunk, int ntx, int nty,
                               int nchunks)
{
    delay_line_body<T, UPWD, ENTROPY, 0, 64, 4, true, false, -1>(a, x, zchunk, ntx, nty, nchunks);
}

// Launch shapes (TX, TY, ZC): 64 x 4 and 64 x 2 run the delay-line kernel above; 32 x 8,
// 16 x 16 and 32 x 16 the ring kernel of solve_ring.hip (32 x 16 in FP32 only: its FP64 ring
// would exceed the 160 KiB LDS).
static const TiledShape kShapes[] = {
    {64, 4, 32}, {64, 4, 64}, {64, 4, 128}, {64, 4, 256},
    {64, 2, 64}, {64, 2, 128},
    {32, 8, 32}, {32, 8, 64}, {32, 8, 128}, {32, 8, 256},
    {16, 16, 32}, {16, 16, 64}, {16, 16, 128}, {16, 16, 256},
    {32, 16, 64}, {32, 16, 128}, {32, 16, 256},
};

static bool is_ring(TiledShape s)
{
    return (s.tx == 32 && s.ty == 8) || (s.tx == 16 && s.ty == 16) || (s.tx == 32 && s.ty == 16);
}

int tiled_num_shapes() { return (int)(sizeof(kShapes) / sizeof(kShapes[0])); }
TiledShape tiled_shape(int i) { return kShapes[i]; }

bool tiled_shape_fits(TiledShape s, int nx, int ny, int nz)
{
    if (s.kind != 0)
        return false;
    bool known = false;
    for (const TiledShape& k : kShapes)
        known = known || (k.tx == s.tx && k.ty == s.ty);
    if (s.tx == 32 && s.ty == 16 && sizeof(AcRealT) != 4)
        return false;
    return known && s.zchunk > 0 && nx % s.tx == 0 && ny % s.ty == 0 && nz % s.zchunk == 0;
}

template <typename T, bool UPWD, bool ENTROPY, int KB, int TX, int TY, bool WRAP>
static void launch_step(const SolveArgs<T>& a, int zchunk, int ntx, int nty, int nch, dim3 bpg,
                        dim3 tpb, hipStream_t s)
{
    if (a.step == 0)
        hipLaunchKernelGGL((solve_tiled_kernel<T, UPWD, ENTROPY, KB, TX, TY, WRAP, 0>), bpg, tpb,
                           0, s, a, zchunk, ntx, nty, nch);
    else if (a.step == 1)
        hipLaunchKernelGGL((solve_tiled_kernel<T, UPWD, ENTROPY, KB, TX, TY, WRAP, 1>), bpg, tpb,
                           0, s, a, zchunk, ntx, nty, nch);
    else
        hipLaunchKernelGGL((solve_tiled_kernel<T, UPWD, ENTROPY, KB, TX, TY, WRAP, 2>), bpg, tpb,
                           0, s, a, zchunk, ntx, nty, nch);
}

template <typename T, bool UPWD, bool ENTROPY, int KB, int TX, int TY>
static void launch_shape(const SolveArgs<T>& a, int zchunk, hipStream_t s)
{
    const int ntx = a.nx / TX, nty = a.ny / TY, nch = a.nz / zchunk;
    const dim3 tpb(TX, TY, 1);
    const dim3 bpg((unsigned)(ntx * nty * nch), 1, 1);
    // FP32: one instantiation per substep, so the substep's branches (w read at s > 0, w written
    // at s < 2, max |u| at s = 2) fold away: 3-11% fewer vector instructions per plane step and
    // no vector spills. FP64 reads the substep at run time (its code is that of the form tested
    // longest; it runs at 512 VGPRs with spills, where code generation has proved fragile).
    if constexpr (sizeof(T) == 4) {
        if (a.wrap)
            launch_step<T, UPWD, ENTROPY, KB, TX, TY, true>(a, zchunk, ntx, nty, nch, bpg, tpb, s);
        else
            launch_step<T, UPWD, ENTROPY, KB, TX, TY, false>(a, zchunk, ntx, nty, nch, bpg, tpb, s);
    }
    else if (a.wrap)
        hipLaunchKernelGGL((solve_tiled_kernel<T, UPWD, ENTROPY, KB, TX, TY, true, -1>), bpg, tpb,
                           0, s, a, zchunk, ntx, nty, nch);
    else
        hipLaunchKernelGGL((solve_tiled_kernel<T, UPWD, ENTROPY, KB, TX, TY, false, -1>), bpg, tpb,
                           0, s, a, zchunk, ntx, nty, nch);
}

template <typename T, bool UPWD, bool ENTROPY, int KB>
static bool launch_kb(const SolveArgs<T>& a, TiledShape sh, hipStream_t s)
{
    if (sh.tx == 64 && sh.ty == 4) launch_shape<T, UPWD, ENTROPY, KB, 64, 4>(a, sh.zchunk, s);
    else if (sh.tx == 64 && sh.ty == 2) launch_shape<T, UPWD, ENTROPY, KB, 64, 2>(a, sh.zchunk, s);
    else return false;
    return true;
}

// The tiled kernels address field f of every role as slab + f * fstride.
template <typename T>
static bool uniform_slabs(const SolveArgs<T>& a, int kb)
{
    for (int q = 0; q < NF; ++q) {
        const long long o = q * a.fstride;
        if (a.in[q] != a.in[0] + o || a.out[q] != a.out[0] + o || a.w[q] != a.w[0] + o)
            return false;
        if (kb > 0 && (const char*)a.comp[q] != (const char*)a.comp[0] + q * (kb / 8))
            return false;
    }
    return true;
}

template <typename T>
bool tiled_reads_periodic(const SolveArgs<T>& a, int kb, TiledShape sh)
{
    return tiled_shape_fits(sh, a.nx, a.ny, a.nz) && uniform_slabs(a, kb) && !is_ring(sh) &&
           a.nx >= R && a.ny >= R && a.nz >= R;
}

template <typename T>
bool launch_solve_tiled(const SolveArgs<T>& a, bool upwd, bool entropy, int kb, TiledShape sh,
                        hipStream_t s)
{
    if (!tiled_shape_fits(sh, a.nx, a.ny, a.nz) || !uniform_slabs(a, kb))
        return false;
    if (a.wrap && !tiled_reads_periodic(a, kb, sh))
        return false;
    if (is_ring(sh))
        return launch_solve_ring<T>(a, upwd, entropy, kb, sh, s);
#define AC_TILED_KB(U, E)                                                                   \
    do {                                                                                    \
        if (kb == 0) return launch_kb<T, U, E, 0>(a, sh, s);                                \
        if (kb == 8) return launch_kb<T, U, E, (sizeof(T) == 4 ? 8 : 0)>(a, sh, s);         \
        if (kb == 16) return launch_kb<T, U, E, (sizeof(T) == 4 ? 16 : 0)>(a, sh, s);       \
        return launch_kb<T, U, E, (sizeof(T) == 4 ? 32 : 0)>(a, sh, s);                     \
    } while (0)
    if (upwd && entropy) AC_TILED_KB(true, true);
    if (upwd) AC_TILED_KB(true, false);
    if (entropy) AC_TILED_KB(false, true);
    AC_TILED_KB(false, false);
#undef AC_TILED_KB
}

template bool launch_solve_tiled<AcRealT>(const SolveArgs<AcRealT>&, bool, bool, int, TiledShape,
                                          hipStream_t);
template bool tiled_reads_periodic<AcRealT>(const SolveArgs<AcRealT>&, int, TiledShape);

template <typename T>
bool launch_adjoint_local_tiled(const SolveArgs<T>& a, const AdjLocal<T>& x, bool upwd,
                                bool entropy, hipStream_t s, long long* nblocks)
{
    if (a.nx % 64 != 0 || a.ny % 4 != 0 || !uniform_slabs(a, 0))
        return false;
    const int zc = adjoint_zchunk(a.nx, a.ny, a.nz);
    const int ntx = a.nx / 64, nty = a.ny / 4, nch = a.nz / zc;
    const dim3 tpb(64, 4, 1), bpg((unsigned)(ntx * nty * nch), 1, 1);
    *nblocks = (long long)ntx * nty * nch;
    if (upwd && entropy) hipLaunchKernelGGL((adjoint_local_tiled_kernel<T, true, true>), bpg, tpb, 0, s, a, x, zc, ntx, nty, nch);
    else if (upwd) hipLaunchKernelGGL((adjoint_local_tiled_kernel<T, true, false>), bpg, tpb, 0, s, a, x, zc, ntx, nty, nch);
    else if (entropy) hipLaunchKernelGGL((adjoint_local_tiled_kernel<T, false, true>), bpg, tpb, 0, s, a, x, zc, ntx, nty, nch);
    else hipLaunchKernelGGL((adjoint_local_tiled_kernel<T, false, false>), bpg, tpb, 0, s, a, x, zc, ntx, nty, nch);
    return true;
}

template bool launch_adjoint_local_tiled<AcRealT>(const SolveArgs<AcRealT>&, const AdjLocal<AcRealT>&,
                                                  bool, bool, hipStream_t, long long*);

}  // namespace acamd
