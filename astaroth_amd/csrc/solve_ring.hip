// solve_ring.hip -- four-plane LDS ring variant of the fused RK3 substep for gfx950.
//
// Same semantics as solve_direct.hip and solve_tiled.hip (ref
// acc-runtime/samples/mhd_modular/mhdsolver.ac:586-636, one fused pass); the difference is where
// the stencil neighbourhood lives while a workgroup marches up its TX x TY column block:
//
//  * The LDS holds a ring of four xy planes (each with its radius-3 halo, all eight fields): the
//    output plane z and the planes z+1 .. z+3 above it.
//  * Each plane step does
//      phase A  in-plane stencils of plane z from its slot: x/y first, second, upwind and
//               xy-mixed derivatives. Plane z's contributions to the backward halves of the
//               xz / yz mixed derivatives of planes z+1 .. z+3 are pushed into registers:
//                   d_xz f(z) = sum_i c_i [Dx_i f(z+i) - Dx_i f(z-i)],  Dx_i f = f(x+i) - f(x-i);
//      barrier  plane z+3 (DMA'd one step earlier) has landed and every wave is done with slot
//               z, which then receives the DMA of plane z+4 (global->LDS, no staging
//               registers) -- it lands while the rest of this step and phase A of the next run;
//      phase B  z-direction stencils: forward halves f(z+1 .. z+3) and the forward halves of
//               the mixed-derivative chains from the ring, backward halves from registers; then
//               the right-hand side and the RK3 update of plane z.
//  * Register state carried between steps is only the column's three previous values per field
//    (24) and three pending backward chain sums for each of the six grad(div) components (18),
//    so the FP32 kernel fits the 256-register budget of two waves per SIMD; with 256-thread
//    workgroups and a ring of 4 x 8 x (TX+6)(TY+6) values (~74 KB at 32 x 8), two workgroups
//    share a CU and hide each other's barrier and LDS latency.
//  * Workgroups are numbered so that consecutive tiles (which share halo rows) land on the same
//    XCD and its L2: hardware dispatches workgroup b to XCD b % 8.
#include "kernels.h"

namespace acamd {

namespace {
constexpr int NSLOT = 4;
constexpr int NCH = 6;  // chains: grad(div u) x, y, z and grad(div A) x, y, z
}  // namespace

template <typename T, bool UPWD, bool ENTROPY, int KB, int TX, int TY>
__global__ void __launch_bounds__(TX* TY, (sizeof(T) == 4 ? 2 : 1))
    solve_ring_kernel(SolveArgs<T> a, int zchunk, int ntx, int nty, int nchunks)
{
    constexpr int LX = TX + 2 * R;
    constexpr int LY = TY + 2 * R;
    constexpr int LP = LX * LY;                  // one field-plane with halo
    constexpr int NT = TX * TY;
    constexpr int NW = NT / 64;                  // waves per workgroup
    constexpr int WPE = (int)sizeof(T) / 4;      // dwords per element
    constexpr int LPW = LP * WPE;                // dwords per field-plane
    constexpr int NCHUNK = (LPW + 63) / 64;      // wave-wide DMA chunks per field-plane
    constexpr int LPS = NCHUNK * 64 / WPE;       // LDS stride of one field-plane (elements)
    constexpr int NR = (NCHUNK + NW - 1) / NW;   // chunks per wave per field-plane (at most)
    constexpr int SLOT = NF * LPS;               // elements per ring slot
    // 32-bit compensation (the precise mode): second, upwind and mixed derivatives from
    // differences to the centre value, exact for neighbouring values, instead of neighbour sums
    constexpr bool CANCEL_FREE = sizeof(T) == 4 && KB == 32;
    static_assert(NT % 64 == 0, "workgroup must be whole waves");
    static_assert((NCHUNK * 64) % WPE == 0, "field-plane stride");
    __shared__ T lds[NSLOT * SLOT];

    // ---- XCD-aware tile numbering (x fastest, then y, then z-chunk per XCD)
    const int nblocks = ntx * nty * nchunks;
    const int bid = blockIdx.x;
    int logical = bid;
    if ((nblocks & 7) == 0)
        logical = (bid & 7) * (nblocks >> 3) + (bid >> 3);
    const int tile_x = logical % ntx;
    const int rest = logical / ntx;
    const int tile_y = rest % nty;
    const int chunk = rest / nty;

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = tx + TX * ty;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x0 = tile_x * TX, y0 = tile_y * TY, z0 = chunk * zchunk;
    const long long sy = a.mx, sz = a.sz;

    // ---- DMA: chunk j = wave + r * NW of a field-plane covers its dwords [64 j, 64 j + 64);
    // lane l moves dword d = 64 j + l. Lanes past the plane's end re-read its first dword and
    // land in the slot's padding, so no lane is masked.
    unsigned g_off[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        int d = (wave + r * NW) * 64 + lane;
        if (d >= LPW)
            d = 0;
        const int e = d / WPE, wd = d - e * WPE;
        const int ly = e / LX, lx = e - ly * LX;
        g_off[r] = (unsigned)(((x0 + lx) + (int)sy * (y0 + ly)) * (int)sizeof(T) + 4 * wd);
    }
    // Field f of a role lives at slab + f * fstride; the per-field bases are recomputed from a
    // stride made opaque once per plane step (scalar work) rather than held across the loop.
    const char* const in0 = (const char*)a.in[0];
    char* const out0 = (char*)a.out[0];
    char* const w0 = (char*)a.w[0];
    char* const comp0 = (char*)a.comp[0];
    auto dma_plane = [&](int p, int slot, long long fsb) {
        const long long base = sz * (long long)(p + R) * (long long)sizeof(T);
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const char* src = in0 + (base + f * fsb);
            char* dst = (char*)&lds[slot * SLOT + f * LPS];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const int j = wave + r * NW;  // wave-uniform: a scalar branch
                if (j < NCHUNK)
                    __builtin_amdgcn_global_load_lds((const void*)(src + opaque(g_off[r])),
                                                     (lds_ptr_t)(dst + 256 * j), 4, 0, 0);
            }
        }
    };

    const Coeffs<T>& C = a.C;
    const T dt = solve_dt(a);
    double u2max = 0.0;  // max |u_new|^2 (the courant timestep mode)
    const int gx = x0 + tx, gy = y0 + ty;
    const int ctr = (ty + R) * LX + (tx + R);
    const unsigned ioff = (unsigned)((gx + R) + (int)sy * (gy + R));  // in-plane offset

    T qb[NF][3];        // f(z-3), f(z-2), f(z-1) of this column
    T bwd[NCH][3];      // backward chain sums pending for planes z, z+1, z+2
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            qb[f][k] = T(0);
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            bwd[c][k] = T(0);

    // prologue: planes z0-3 .. z0 into slots 0 .. 3
#pragma unroll
    for (int k = 0; k < NSLOT; ++k)
        dma_plane(z0 - R + k, k, a.fstride * (long long)sizeof(T));
    plane_barrier();

    const int niter = zchunk + R;
    for (int it = 0; it < niter; ++it) {
        const int z = z0 - R + it;     // plane of phase A (and the output plane once z >= z0)
        const int s0 = it & 3;
        const bool out = it >= R;
        const unsigned iob = opaque(ioff * (unsigned)sizeof(T));  // byte offset in a plane
        const long long fsb = opaque_s64(a.fstride * (long long)sizeof(T));  // field stride, bytes

        // w (and the Kahan compensation) of the output cell, loaded early; consumed at the end
        T wpre[NF];
        CompRec<KB> crec{};  // the output cell's compensation record (KB > 0)
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            wpre[f] = T(0);
        }
        if (out) {
            const long long poff = sz * (long long)(z + R);
            if (a.step > 0) {
#pragma unroll
                for (int f = 0; f < NF; ++f)
                    wpre[f] = ld_b((const T*)(w0 + (f * fsb + poff * (long long)sizeof(T))), iob);
            }
            if (KB > 0)
                crec = comp_rec_load<KB>(comp0 + poff * (long long)(NF * KB / 8),
                                         iob / (unsigned)sizeof(T) * (unsigned)(NF * KB / 8));
        }

        // ================= phase A: in-plane stencils of plane z (slot s0)
        const T* P0 = &lds[s0 * SLOT + ctr];
        auto at0 = [&](int f, int dx, int dy) -> T { return P0[f * LPS + dy * LX + dx]; };
        T cen[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f)
            cen[f] = at0(f, 0, 0);
        const T aux = ac_abs(cen[1]), auy = ac_abs(cen[2]);

        T lr_dx, lr_dy, lr_lap, lr_up;
        T u_dx[3], u_dy[3], u_lap[3], u_up[3];
        T ax_dy, ay_dx, az_dx, az_dy, a_lap[3];
        T s_dx, s_dy, s_lap;
        T gip[NCH] = {};  // in-plane parts of grad(div u), grad(div A) (z components stay 0)
        // Backward halves of the xz / yz chains: plane z contributes -c_i D_i to plane z+i
        // (chains 0/3: d_xz v_z; 1/4: d_yz v_z; 2/5: d_xz v_x + d_yz v_y). Planes z+1 and z+2
        // have pending sums in bwd[.][1..2]; plane z+3 starts its sums in nb3.
        T nb3[NCH] = {};
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const T c0 = cen[f];
            T Dx[4], Dy[4], Sx[4], Sy[4];
#pragma unroll
            for (int i = 1; i <= 3; ++i) {
                const T xp = at0(f, i, 0), xm = at0(f, -i, 0);
                const T yp = at0(f, 0, i), ym = at0(f, 0, -i);
                Dx[i] = xp - xm;
                Sx[i] = CANCEL_FREE ? (xp - c0) + (xm - c0) : xp + xm;
                Dy[i] = yp - ym;
                Sy[i] = CANCEL_FREE ? (yp - c0) + (ym - c0) : yp + ym;
            }
            if (f >= 1 && f <= 6) {
                const int v = f < 4 ? 0 : 1;
                const int fv = f - 1 - 3 * v;  // component 0, 1, 2 of u or A
#pragma unroll
                for (int i = 1; i <= 3; ++i) {
                    // chain 3v+2 gets c_xz Dx v_x and c_yz Dy v_y; 3v+0 / 3v+1 get Dx / Dy v_z
                    T e0 = T(0), e1 = T(0), e2 = T(0);
                    if (fv == 0) e2 = C.mx[1][i] * Dx[i];
                    if (fv == 1) e2 = C.mx[2][i] * Dy[i];
                    if (fv == 2) {
                        e0 = C.mx[1][i] * Dx[i];
                        e1 = C.mx[2][i] * Dy[i];
                    }
                    if (i < 3) {
                        if (fv == 2) {
                            bwd[3 * v + 0][i] -= e0;
                            bwd[3 * v + 1][i] -= e1;
                        }
                        else {
                            bwd[3 * v + 2][i] -= e2;
                        }
                    }
                    else {
                        if (fv == 2) {
                            nb3[3 * v + 0] = -e0;
                            nb3[3 * v + 1] = -e1;
                        }
                        else {
                            nb3[3 * v + 2] -= e2;
                        }
                    }
                }
            }
            // the centre weight drops out of the cancellation-free form (the weights sum to zero)
            auto sym = [&](const T* k, const T* S2) -> T {
                if constexpr (CANCEL_FREE)
                    return k[1] * S2[1] + k[2] * S2[2] + k[3] * S2[3];
                else
                    return k[0] * c0 + k[1] * S2[1] + k[2] * S2[2] + k[3] * S2[3];
            };
            const T dxx = sym(C.d2[0], Sx);
            const T dyy = sym(C.d2[1], Sy);
            T upxy = T(0);
            if (UPWD && f < 4) {
                const T upx = sym(C.up[0], Sx);
                const T upy = sym(C.up[1], Sy);
                upxy = aux * upx + auy * upy;
            }
            const bool need_dx = (f != 4);
            const bool need_dy = (f != 5);
            T dx = T(0), dy = T(0);
            if (need_dx)
                dx = C.d1[0][1] * Dx[1] + C.d1[0][2] * Dx[2] + C.d1[0][3] * Dx[3];
            if (need_dy)
                dy = C.d1[1][1] * Dy[1] + C.d1[1][2] * Dy[2] + C.d1[1][3] * Dy[3];
            const T lap = dxx + dyy;
            if (f == 0) {
                lr_dx = dx;
                lr_dy = dy;
                lr_lap = lap;
                lr_up = upxy;
            }
            else if (f < 4) {
                u_dx[f - 1] = dx;
                u_dy[f - 1] = dy;
                u_lap[f - 1] = lap;
                u_up[f - 1] = upxy;
            }
            else if (f == 4) {
                ax_dy = dy;
                a_lap[0] = lap;
            }
            else if (f == 5) {
                ay_dx = dx;
                a_lap[1] = lap;
            }
            else if (f == 6) {
                az_dx = dx;
                az_dy = dy;
                a_lap[2] = lap;
            }
            else {
                s_dx = dx;
                s_dy = dy;
                s_lap = lap;
            }
            // in-plane second derivatives of grad(div v): x gets d_xx v_x, y gets d_yy v_y
            if (f == 1) gip[0] = dxx;
            if (f == 2) gip[1] = dyy;
            if (f == 4) gip[3] = dxx;
            if (f == 5) gip[4] = dyy;
            // one field's neighbourhood in flight at a time keeps the register peak low
            __builtin_amdgcn_sched_barrier(0);
        }
        // xy-mixed derivatives: d_xy v_y -> (grad div v)_x, d_xy v_x -> (grad div v)_y
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const int fx = 1 + 3 * v, fy = 2 + 3 * v;
            T mxx = T(0), myy = T(0);
#pragma unroll
            for (int i = 1; i <= 3; ++i) {
                if constexpr (CANCEL_FREE) {
                    mxx += C.mx[0][i] * ((at0(fx, i, i) - at0(fx, i, -i)) - (at0(fx, -i, i) - at0(fx, -i, -i)));
                    myy += C.mx[0][i] * ((at0(fy, i, i) - at0(fy, i, -i)) - (at0(fy, -i, i) - at0(fy, -i, -i)));
                }
                else {
                    mxx += C.mx[0][i] * ((at0(fx, i, i) + at0(fx, -i, -i)) - (at0(fx, i, -i) + at0(fx, -i, i)));
                    myy += C.mx[0][i] * ((at0(fy, i, i) + at0(fy, -i, -i)) - (at0(fy, i, -i) + at0(fy, -i, i)));
                }
            }
            gip[3 * v + 0] += myy;
            gip[3 * v + 1] += mxx;
            __builtin_amdgcn_sched_barrier(0);
        }
        __builtin_amdgcn_sched_barrier(0);

        // ================= plane z+3 has landed everywhere; slot s0 is free for plane z+4
        plane_barrier();
        if (it + 1 < niter)
            dma_plane(z + 4, s0, fsb);
        __builtin_amdgcn_sched_barrier(0);

        // ================= phase B: z stencils, chains, right-hand side, RK3 update
        if (out) {
            const T* P1 = &lds[((it + 1) & 3) * SLOT + ctr];
            const T* P2 = &lds[((it + 2) & 3) * SLOT + ctr];
            const T* P3 = &lds[((it + 3) & 3) * SLOT + ctr];
            auto atk = [&](int k, int f, int dx, int dy) -> T {
                const T* Pk = k == 1 ? P1 : (k == 2 ? P2 : P3);
                return Pk[f * LPS + dy * LX + dx];
            };
            T dz[NF], dzz[NF], upz[4];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const T fp[4] = {cen[f], atk(1, f, 0, 0), atk(2, f, 0, 0), atk(3, f, 0, 0)};
                T D[4], S[4];
#pragma unroll
                for (int i = 1; i <= 3; ++i) {
                    D[i] = fp[i] - qb[f][3 - i];
                    S[i] = CANCEL_FREE ? (fp[i] - cen[f]) + (qb[f][3 - i] - cen[f])
                                       : fp[i] + qb[f][3 - i];
                }
                dz[f] = C.d1[2][1] * D[1] + C.d1[2][2] * D[2] + C.d1[2][3] * D[3];
                if constexpr (CANCEL_FREE) {
                    dzz[f] = C.d2[2][1] * S[1] + C.d2[2][2] * S[2] + C.d2[2][3] * S[3];
                    if (UPWD && f < 4)
                        upz[f] = C.up[2][1] * S[1] + C.up[2][2] * S[2] + C.up[2][3] * S[3];
                }
                else {
                    dzz[f] = C.d2[2][0] * cen[f] + C.d2[2][1] * S[1] + C.d2[2][2] * S[2] + C.d2[2][3] * S[3];
                    if (UPWD && f < 4)
                        upz[f] = C.up[2][0] * cen[f] + C.up[2][1] * S[1] + C.up[2][2] * S[2] + C.up[2][3] * S[3];
                }
                if (f & 1)
                    __builtin_amdgcn_sched_barrier(0);
            }
            // forward halves of the chains: + c_i D_i at planes z+i
            T fwd[NCH];
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const int fx = 1 + 3 * v, fy = 2 + 3 * v, fz = 3 + 3 * v;
                T ex = T(0), ey = T(0), ez = T(0);
#pragma unroll
                for (int i = 1; i <= 3; ++i) {
                    ex += C.mx[1][i] * (atk(i, fz, i, 0) - atk(i, fz, -i, 0));
                    ey += C.mx[2][i] * (atk(i, fz, 0, i) - atk(i, fz, 0, -i));
                    ez += C.mx[1][i] * (atk(i, fx, i, 0) - atk(i, fx, -i, 0)) +
                          C.mx[2][i] * (atk(i, fy, 0, i) - atk(i, fy, 0, -i));
                }
                fwd[3 * v + 0] = ex;
                fwd[3 * v + 1] = ey;
                fwd[3 * v + 2] = ez;
                __builtin_amdgcn_sched_barrier(0);
            }

            CellDerivs<T> d;
            d.lr = cen[0];
            d.glr[0] = lr_dx;
            d.glr[1] = lr_dy;
            d.glr[2] = dz[0];
            d.laplr = lr_lap + dzz[0];
            const T auz = ac_abs(cen[3]);
            d.upl = UPWD ? lr_up + auz * upz[0] : T(0);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                d.u[i] = cen[1 + i];
                d.gu[i][0] = u_dx[i];
                d.gu[i][1] = u_dy[i];
                d.gu[i][2] = dz[1 + i];
                d.lapu[i] = u_lap[i] + dzz[1 + i];
                d.upu[i] = UPWD ? u_up[i] + auz * upz[1 + i] : T(0);
                d.lapa[i] = a_lap[i] + dzz[4 + i];
            }
            d.gdu[0] = gip[0] + (fwd[0] + bwd[0][0]);
            d.gdu[1] = gip[1] + (fwd[1] + bwd[1][0]);
            d.gdu[2] = (fwd[2] + bwd[2][0]) + dzz[3];
            d.gda[0] = gip[3] + (fwd[3] + bwd[3][0]);
            d.gda[1] = gip[4] + (fwd[4] + bwd[4][0]);
            d.gda[2] = (fwd[5] + bwd[5][0]) + dzz[6];
            d.ga[0][0] = T(0);
            d.ga[0][1] = ax_dy;
            d.ga[0][2] = dz[4];
            d.ga[1][0] = ay_dx;
            d.ga[1][1] = T(0);
            d.ga[1][2] = dz[5];
            d.ga[2][0] = az_dx;
            d.ga[2][1] = az_dy;
            d.ga[2][2] = T(0);
            d.s = cen[7];
            d.gs[0] = s_dx;
            d.gs[1] = s_dy;
            d.gs[2] = dz[7];
            d.laps = s_lap + dzz[7];

            T rhs[NF];
            mhd_rhs<T, UPWD, ENTROPY, CANCEL_FREE>(d, a.P, rhs);
            const long long poff = sz * (long long)(z + R);
            T fnew[NF];
            CompRec<KB> cout = crec;  // ss keeps its record entry without entropy
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const long long fo = f * fsb + poff * (long long)sizeof(T);  // bytes
                T* op = (T*)(out0 + fo);
                if (f == 7 && !ENTROPY) {
                    fnew[f] = cen[f];
                    st_b(op, iob, cen[f]);
                    continue;
                }
                T* wp = (T*)(w0 + fo);
                int cenc = 0;
                fnew[f] = rk_update_pre<T, KB>(wp, op, iob, a.step, a.alpha, a.beta, cen[f], rhs[f],
                                               dt, wpre[f], comp_rec_get<KB>(crec, f), &cenc);
                if (KB > 0)
                    comp_rec_set<KB>(cout, f, cenc);
            }
            if (KB > 0)
                comp_rec_store<KB>(comp0 + poff * (long long)(NF * KB / 8),
                                   iob / (unsigned)sizeof(T) * (unsigned)(NF * KB / 8), cout);
            if (a.step == 2) {
                const double u2 = (double)fnew[1] * (double)fnew[1] + (double)fnew[2] * (double)fnew[2] +
                                  (double)fnew[3] * (double)fnew[3];
                u2max = fmax(u2max, u2);
            }
        }

        // ================= slide the register window by one plane
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            qb[f][0] = qb[f][1];
            qb[f][1] = qb[f][2];
            qb[f][2] = cen[f];
        }
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            bwd[c][0] = bwd[c][1];
            bwd[c][1] = bwd[c][2];
            bwd[c][2] = nb3[c];
        }
    }
    if (a.step == 2) {
        if (a.umax2_next != nullptr)
            wave_max_to_global(u2max, a.umax2_next);
        if (a.time_acc != nullptr && blockIdx.x == 0 && tid == 0)
            *a.time_acc += (double)dt;
    }
}

template <typename T, bool UPWD, bool ENTROPY, int KB, int TX, int TY>
static void launch_ring(const SolveArgs<T>& a, int zchunk, hipStream_t s)
{
    const int ntx = a.nx / TX, nty = a.ny / TY, nch = a.nz / zchunk;
    const dim3 tpb(TX, TY, 1);
    const dim3 bpg((unsigned)(ntx * nty * nch), 1, 1);
    hipLaunchKernelGGL((solve_ring_kernel<T, UPWD, ENTROPY, KB, TX, TY>), bpg, tpb, 0, s, a,
                       zchunk, ntx, nty, nch);
}

template <typename T, bool UPWD, bool ENTROPY, int KB>
static bool ring_kb(const SolveArgs<T>& a, TiledShape sh, hipStream_t s)
{
    if (sh.tx == 32 && sh.ty == 8) launch_ring<T, UPWD, ENTROPY, KB, 32, 8>(a, sh.zchunk, s);
    else if (sh.tx == 16 && sh.ty == 16) launch_ring<T, UPWD, ENTROPY, KB, 16, 16>(a, sh.zchunk, s);
    else if (sh.tx == 32 && sh.ty == 16) {
        if constexpr (sizeof(T) == 4)
            launch_ring<T, UPWD, ENTROPY, KB, 32, 16>(a, sh.zchunk, s);
        else
            return false;
    }
    else return false;
    return true;
}

template <typename T>
bool launch_solve_ring(const SolveArgs<T>& a, bool upwd, bool entropy, int kb, TiledShape sh,
                       hipStream_t s)
{
#define AC_RING_KB(U, E)                                                                    \
    do {                                                                                    \
        if (kb == 0) return ring_kb<T, U, E, 0>(a, sh, s);                                  \
        if (kb == 8) return ring_kb<T, U, E, (sizeof(T) == 4 ? 8 : 0)>(a, sh, s);           \
        if (kb == 16) return ring_kb<T, U, E, (sizeof(T) == 4 ? 16 : 0)>(a, sh, s);         \
        return ring_kb<T, U, E, (sizeof(T) == 4 ? 32 : 0)>(a, sh, s);                       \
    } while (0)
    if (upwd && entropy) AC_RING_KB(true, true);
    if (upwd) AC_RING_KB(true, false);
    if (entropy) AC_RING_KB(false, true);
    AC_RING_KB(false, false);
#undef AC_RING_KB
}

template bool launch_solve_ring<AcRealT>(const SolveArgs<AcRealT>&, bool, bool, int, TiledShape,
                                         hipStream_t);

}  // namespace acamd
