// om_closure_tiled.hip — z-marching, LDS-tiled eddy viscosity / diffusivity kernel
// (AnisotropicMinimumDissipation: anisotropic_minimum_dissipation.jl:161-358; SmagorinskyLilly:
// smagorinsky.jl:90-104, lilly_coefficient.jl:129-142) for three-dimensional grids with Periodic x, y.
//
// A block of 16 × 16 threads owns 16 × 16 columns and marches up a chunk of z levels.  The planes
// u and v (levels k-1, k, k+1), w (k, k+1) and the tracers (k), each with a one-cell ring, are
// staged in LDS; every velocity gradient of a cell's AMD composite is then read from LDS instead
// of from L1/L2.  Each thread keeps its column's tracers at k-1 .. k+1 in registers for the
// vertical gradients.  The expressions and their operand order are those of k_closure
// (om_kernels.hip), so the two kernels give identically rounded results.
#include <hip/hip_runtime.h>

#include "om_batch.h"
#include "om_kernels.h"

namespace om {

constexpr int CT = 16;              // cells (and threads) per tile edge
constexpr int CW = CT + 2;          // staged plane width: i0-1 .. i0+16
constexpr int CA = CW * CW;

template <class T, int NT> struct ClosureSmem {
    T u[3][CA], v[3][CA], w[2][CA];
    T c[NT > 0 ? NT : 1][CA];
};

template <class T> __device__ __forceinline__ void stage_ring(T *dst, const T *f, const DevGrid<T> &g, int i0, int j0, int k,
                                                              int tid) {
    const int imax = g.Nx + g.Hx - 1, jmax = g.Ny + g.Hy - 1;
    for (int e = tid; e < CA; e += CT * CT) {
        const int y = e / CW, x = e - y * CW;
        const int i = i0 - 1 + x, j = j0 - 1 + y;
        dst[e] = (i <= imax && j <= jmax) ? f[idx3(i, j, k, g.sy, g.sz)] : T(0);
    }
}

template <class T> struct ClosureTiledArgs {
    Ctx<T> C;
    int kchunk, ntx;
};

template <class T, int NT> struct KClosureTiled {
    static constexpr int kThreads = 256, kMinBlocks = 1;
    static __device__ __forceinline__ void run(const ClosureTiledArgs<T> &A, const Blk &B);
};

template <class T, int NT>
__device__ __forceinline__ void KClosureTiled<T, NT>::run(const ClosureTiledArgs<T> &A, const Blk &B) {
    __shared__ ClosureSmem<T, NT> S;
    const Ctx<T> &C = A.C;
    const int kchunk = A.kchunk, ntx = A.ntx;
    const DevGrid<T> &g = C.g;
    const DevPhysics<T> &ph = C.ph;
    const int tid = threadIdx.x;
    const int a = tid & (CT - 1), b = tid >> 4;
    // work item → (z chunk, tile): the chunks of one tile are consecutive, tiles in row order
    const int item = xcd_work_item((int)B.x, (int)B.gx);
    const int nch = (g.Nz + kchunk - 1) / kchunk;
    const int tile = item / nch, chunk = item - tile * nch;
    const int tx = tile % ntx, ty = tile / ntx;
    const int i0 = tx * CT, j0 = ty * CT;
    const int k0 = chunk * kchunk;
    const int k1 = min(k0 + kchunk, g.Nz);
    const int i = i0 + a, j = j0 + b;
    const bool colok = i <= g.Nx + g.Hx - 1 && j <= g.Ny + g.Hy - 1;
    const bool cell = i < g.Nx && j < g.Ny;
    const int64_t col = idx3(i, j, 0, g.sy, g.sz);
    const bool smag = ph.closure == OM_CLOSURE_SMAGORINSKY_LILLY;

    auto P = [&](const T *p, int dx, int dy) -> T { return p[(b + 1 + dy) * CW + (a + 1 + dx)]; };
    auto slot3 = [](int level) { return ((level % 3) + 3) % 3; };
    // own-column access through scalar level origins and one per-thread byte offset (the column's
    // offset within a level, < 2³² bytes), refreshed opaquely each level so that no per-field
    // pointer stays live across the level's arithmetic
    const uint32_t colb0 = (uint32_t)(col * (int64_t)sizeof(T));
    uint32_t colb = colb0;
    auto ldcol = [&](const T *f, int k) -> T {
        return colok ? *reinterpret_cast<const T *>(reinterpret_cast<const char *>(f + (int64_t)k * g.sz) + colb) : T(0);
    };

    // own-column tracers at levels kk-1, kk, kk+1 ([0], [1], [2])
    T cw[NT > 0 ? NT : 1][3];
#pragma unroll
    for (int n = 0; n < NT; ++n) { cw[n][1] = ldcol(C.s.U[3 + n], k0 - 1); cw[n][2] = ldcol(C.s.U[3 + n], k0); }
    stage_ring(S.u[slot3(k0 - 1)], C.s.U[0], g, i0, j0, k0 - 1, tid);
    stage_ring(S.v[slot3(k0 - 1)], C.s.U[1], g, i0, j0, k0 - 1, tid);
    stage_ring(S.u[slot3(k0)], C.s.U[0], g, i0, j0, k0, tid);
    stage_ring(S.v[slot3(k0)], C.s.U[1], g, i0, j0, k0, tid);
    stage_ring(S.w[k0 & 1], C.s.U[2], g, i0, j0, k0, tid);

    const T half = T(0.5);
    // The planes a level adds (u, v, w at k + 1, the tracers at k) are loaded into registers while
    // the previous level is evaluated, so their memory latency is hidden behind that work; the
    // plane shares are rebuilt for each load from an opaque copy of tid.
    T ru[2], rv[2], rw[2], rc[NT > 0 ? NT : 1][2];
    auto load_level = [&](int k) {
        int otid = tid;
        asm volatile("" : "+v"(otid));
        const PlaneShare<T, CW> ps(g, i0, j0, 1, otid);
        ps.load(ru, ps.origin(C.s.U[0], g, i0, j0, 1, k + 1));
        ps.load(rv, ps.origin(C.s.U[1], g, i0, j0, 1, k + 1));
        ps.load(rw, ps.origin(C.s.U[2], g, i0, j0, 1, k + 1));
#pragma unroll
        for (int n = 0; n < NT; ++n) ps.load(rc[n], ps.origin(C.s.U[3 + n], g, i0, j0, 1, k));
    };
    load_level(k0);
    // AMD: the x-z and y-z gradients at level k + 1 of one level are those at level k of the next
    // (same expressions on the same planes and ratios), so they are carried instead of recomputed
    T c_xw[2], c_zu[2], c_ywx[2], c_yw[2], c_zv[2];
    bool carried = false;
    for (int k = k0; k < k1; ++k) {
        int otid = tid;
        asm volatile("" : "+v"(otid));
        colb = colb0;
        asm volatile("" : "+v"(colb));
        const PlaneShare<T, CW> ps(g, i0, j0, 1, otid);
#pragma unroll
        for (int n = 0; n < NT; ++n) { cw[n][0] = cw[n][1]; cw[n][1] = cw[n][2]; cw[n][2] = ldcol(C.s.U[3 + n], k + 1); }
        __syncthreads();   // previous level's readers are done
        ps.store(S.u[slot3(k + 1)], ru, tid);
        ps.store(S.v[slot3(k + 1)], rv, tid);
        ps.store(S.w[(k + 1) & 1], rw, tid);
#pragma unroll
        for (int n = 0; n < NT; ++n) ps.store(S.c[n], rc[n], tid);
        __syncthreads();
        if (k + 1 < k1) load_level(k + 1);
        if (!cell) continue;

        const T *Um = S.u[slot3(k - 1)], *U0 = S.u[slot3(k)], *Up = S.u[slot3(k + 1)];
        const T *Vm = S.v[slot3(k - 1)], *V0 = S.v[slot3(k)], *Vp = S.v[slot3(k + 1)];
        const T *W0 = S.w[k & 1], *W1 = S.w[(k + 1) & 1];
        const T idx = g.idx, idy = g.idy;
        // gradients at the points named by the Ctx operators (velocity_tracer_gradients.jl:6-18)
        const T dxu = (P(U0, 1, 0) - P(U0, 0, 0)) * idx;
        const T dyv = (P(V0, 0, 1) - P(V0, 0, 0)) * idy;
        const T dzw = (P(W1, 0, 0) - P(W0, 0, 0)) * g.idzc[k];
        auto dx_v = [&](int di, int dj) { return (P(V0, di, dj) - P(V0, di - 1, dj)) * idx; };          // ffc (i+di, j+dj, k)
        auto dy_u = [&](int di, int dj) { return (P(U0, di, dj) - P(U0, di, dj - 1)) * idy; };          // ffc
        auto dx_w = [&](const T *Wl, int di) { return (P(Wl, di, 0) - P(Wl, di - 1, 0)) * idx; };      // fcf (i+di, j, kk)
        auto dz_u = [&](const T *Ul, const T *Ulm, int di, int kk) { return (P(Ul, di, 0) - P(Ulm, di, 0)) * g.idzf[kk]; };
        auto dy_w = [&](const T *Wl, int di, int dj) { return (P(Wl, di, dj) - P(Wl, di, dj - 1)) * idy; };
        auto dz_v = [&](const T *Vl, const T *Vlm, int dj, int kk) { return (P(Vl, 0, dj) - P(Vlm, 0, dj)) * g.idzf[kk]; };
        auto at = [&](T *f) -> T & { return *reinterpret_cast<T *>(reinterpret_cast<char *>(f + (int64_t)k * g.sz) + colb); };

        if (smag) {
            auto S12 = [&](int di, int dj) { return half * (dy_u(di, dj) + dx_v(di, dj)); };
            T s12[4] = {S12(0, 0), S12(1, 0), S12(0, 1), S12(1, 1)};
            T s13[4] = {half * (dz_u(U0, Um, 0, k) + dx_w(W0, 0)), half * (dz_u(U0, Um, 1, k) + dx_w(W0, 1)),
                        half * (dz_u(Up, U0, 0, k + 1) + dx_w(W1, 0)), half * (dz_u(Up, U0, 1, k + 1) + dx_w(W1, 1))};
            T s23[4] = {half * (dz_v(V0, Vm, 0, k) + dy_w(W0, 0, 0)), half * (dz_v(V0, Vm, 1, k) + dy_w(W0, 0, 1)),
                        half * (dz_v(Vp, V0, 0, k + 1) + dy_w(W1, 0, 0)), half * (dz_v(Vp, V0, 1, k + 1) + dy_w(W1, 0, 1))};
            T tr = dxu * dxu + dyv * dyv + dzw * dzw;
            T S2 = tr + T(2) * avg4<T>(s12[0] * s12[0], s12[1] * s12[1], s12[2] * s12[2], s12[3] * s12[3])
                      + T(2) * avg4<T>(s13[0] * s13[0], s13[1] * s13[1], s13[2] * s13[2], s13[3] * s13[3])
                      + T(2) * avg4<T>(s23[0] * s23[0], s23[1] * s23[1], s23[2] * s23[2], s23[3] * s23[3]);
            T D3 = (g.dx * g.dy) * g.dzc[k];
            T Df = cbrt(D3);
            // ∂z b at ccf k and k+1 from this column's tracers
            auto dzb = [&](int lo, int kk) -> T {   // (c[lo+1] - c[lo]) · idzf[kk]
                if (ph.buoyancy == OM_BUOYANCY_TRACER) {
                    T r = T(0);
#pragma unroll
                    for (int n = 0; n < NT; ++n) if (n == ph.b_tracer) r = (cw[n][lo + 1] - cw[n][lo]) * g.idzf[kk];
                    return r;
                }
                if (ph.buoyancy == OM_BUOYANCY_LINEAR_EOS) {
                    T dT = T(0), dS = T(0);
#pragma unroll
                    for (int n = 0; n < NT; ++n) {
                        if (n == ph.T_tracer) dT = (cw[n][lo + 1] - cw[n][lo]) * g.idzf[kk];
                        if (n == ph.S_tracer) dS = (cw[n][lo + 1] - cw[n][lo]) * g.idzf[kk];
                    }
                    return ph.g * (ph.alpha * dT - ph.beta * dS);
                }
                return T(0);
            };
            T N2 = half * (dzb(0, k) + dzb(1, k + 1));
            T c0 = ph.c0, cb = ph.c1;
            T N2p = fmax(T(0), N2);
            T s2 = T(1) - fmin(T(1), cb * N2p / S2);
            T sig = (S2 == T(0)) ? T(0) : sqrt(s2);
            at(C.s.nu) = (sig * (c0 * c0)) * (Df * Df) * sqrt(T(2) * S2);
            continue;
        }

        // ---- AMD (the level ratios come from the grid's table: same values, no per-thread division)
        const T Dfx = T(2) * g.dx, Dfy = T(2) * g.dy;
        const T Dfz0 = T(2) * g.dzc[k], Dfz1 = T(2) * g.dzc[k + 1];
        const T *r0 = g.amdr + (size_t)(k + 1) * 5, *r1 = r0 + 5;
        const T rxz0 = r0[0], rzx0 = r0[1], ryz0 = r0[2], rzy0 = r0[3];
        const T rxz1 = r1[0], rzx1 = r1[1], ryz1 = r1[2], rzy1 = r1[3];
        T ndxv[4], ndyu[4];
        {
            const int di[4] = {0, 1, 0, 1}, dj[4] = {0, 0, 1, 1};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                ndxv[q] = Dfx / Dfy * dx_v(di[q], dj[q]);
                ndyu[q] = Dfy / Dfx * dy_u(di[q], dj[q]);
            }
        }
        T ndxw[4], ndzu[4], ndyw_x[4];
        T ndyw[4], ndzv[4];
        // x-z points (i,k), (i+1,k), (i,k+1), (i+1,k+1); y-z points (j,k), (j+1,k), (j,k+1), (j+1,k+1)
        if (carried) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                ndxw[q] = c_xw[q]; ndzu[q] = c_zu[q]; ndyw_x[q] = c_ywx[q]; ndyw[q] = c_yw[q]; ndzv[q] = c_zv[q];
            }
        } else {
            ndxw[0] = rxz0 * dx_w(W0, 0);            ndzu[0] = rzx0 * dz_u(U0, Um, 0, k);
            ndxw[1] = rxz0 * dx_w(W0, 1);            ndzu[1] = rzx0 * dz_u(U0, Um, 1, k);
            ndyw_x[0] = ryz0 * dy_w(W0, 0, 0);
            ndyw_x[1] = ryz0 * dy_w(W0, 1, 0);
            ndyw[0] = ryz0 * dy_w(W0, 0, 0);         ndzv[0] = rzy0 * dz_v(V0, Vm, 0, k);
            ndyw[1] = ryz0 * dy_w(W0, 0, 1);         ndzv[1] = rzy0 * dz_v(V0, Vm, 1, k);
        }
        ndxw[2] = rxz1 * dx_w(W1, 0);            ndzu[2] = rzx1 * dz_u(Up, U0, 0, k + 1);
        ndxw[3] = rxz1 * dx_w(W1, 1);            ndzu[3] = rzx1 * dz_u(Up, U0, 1, k + 1);
        ndyw_x[2] = ryz1 * dy_w(W1, 0, 0);
        ndyw_x[3] = ryz1 * dy_w(W1, 1, 0);
        ndyw[2] = ryz1 * dy_w(W1, 0, 0);         ndzv[2] = rzy1 * dz_v(Vp, V0, 0, k + 1);
        ndyw[3] = ryz1 * dy_w(W1, 0, 1);         ndzv[3] = rzy1 * dz_v(Vp, V0, 1, k + 1);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            c_xw[q] = ndxw[2 + q]; c_zu[q] = ndzu[2 + q]; c_ywx[q] = ndyw_x[2 + q]; c_yw[q] = ndyw[2 + q]; c_zv[q] = ndzv[2 + q];
        }
        carried = true;
        T nS12[4], nS13[4], nS23[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            nS12[q] = half * (ndyu[q] + ndxv[q]);
            nS13[q] = half * (ndzu[q] + ndxw[q]);
            nS23[q] = half * (ndzv[q] + ndyw[q]);
        }
#define A4(expr) avg4<T>([&](int q) { return expr; }(0), [&](int q) { return expr; }(1), [&](int q) { return expr; }(2), [&](int q) { return expr; }(3))
        const T q2 = dxu * dxu + dyv * dyv + dzw * dzw
                   + A4(ndxv[q] * ndxv[q]) + A4(ndyu[q] * ndyu[q])
                   + A4(ndxw[q] * ndxw[q]) + A4(ndzu[q] * ndzu[q])
                   + A4(ndyw[q] * ndyw[q]) + A4(ndzv[q] * ndzv[q]);
        const T delta2 = r0[4];
        T nu = T(0);
        if (q2 != T(0)) {
            const T S11 = dxu, S22 = dyv, S33 = dzw;
            const T t1 = S11 * (dxu * dxu) + S22 * A4(ndxv[q] * ndxv[q]) + S33 * A4(ndxw[q] * ndxw[q])
                       + T(2) * dxu * A4(ndxv[q] * nS12[q]) + T(2) * dxu * A4(ndxw[q] * nS13[q])
                       + T(2) * A4(ndxv[q]) * A4(ndxw[q]) * A4(nS23[q]);
            const T t2 = S11 * A4(ndyu[q] * ndyu[q]) + S22 * (dyv * dyv) + S33 * A4(ndyw[q] * ndyw[q])
                       + T(2) * dyv * A4(ndyu[q] * nS12[q]) + T(2) * A4(ndyu[q]) * A4(ndyw[q]) * A4(nS13[q])
                       + T(2) * dyv * A4(ndyw[q] * nS23[q]);
            const T t3 = S11 * A4(ndzu[q] * ndzu[q]) + S22 * A4(ndzv[q] * ndzv[q]) + S33 * (dzw * dzw)
                       + T(2) * A4(ndzu[q]) * A4(ndzv[q]) * A4(nS12[q])
                       + T(2) * dzw * A4(ndzu[q] * nS13[q]) + T(2) * dzw * A4(ndzv[q] * nS23[q]);
            const T r = (t1 + t2) + t3;
            nu = ft_div<T>(-ph.c0 * delta2 * (r - T(0)), q2);
        }
        at(C.s.nu) = fmax(T(0), nu);

#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const T *Cn = S.c[n];
            const T gx0 = Dfx * ((P(Cn, 0, 0) - P(Cn, -1, 0)) * idx), gx1 = Dfx * ((P(Cn, 1, 0) - P(Cn, 0, 0)) * idx);
            const T gy0 = Dfy * ((P(Cn, 0, 0) - P(Cn, 0, -1)) * idy), gy1 = Dfy * ((P(Cn, 0, 1) - P(Cn, 0, 0)) * idy);
            const T gz0 = Dfz0 * ((cw[n][1] - cw[n][0]) * g.idzf[k]), gz1 = Dfz1 * ((cw[n][2] - cw[n][1]) * g.idzf[k + 1]);
            const T Ix2 = half * (gx0 * gx0 + gx1 * gx1), Iy2 = half * (gy0 * gy0 + gy1 * gy1), Iz2 = half * (gz0 * gz0 + gz1 * gz1);
            const T sig = (Ix2 + Iy2) + Iz2;
            T kap = T(0);
            if (sig != T(0)) {
                const T Ix = half * (gx0 + gx1), Iy = half * (gy0 + gy1), Iz = half * (gz0 + gz1);
                const T cx_ux = dxu * Ix2 + A4(ndxv[q]) * Ix * Iy + A4(ndxw[q]) * Ix * Iz;
                const T cy_uy = A4(ndyu[q]) * Iy * Ix + dyv * Iy2 + A4(ndyw_x[q]) * Iy * Iz;
                const T cz_uz = A4(ndzu[q]) * Iz * Ix + A4(ndzv[q]) * Iz * Iy + dzw * Iz2;
                const T theta = (cx_ux + cy_uy) + cz_uz;
                kap = ft_div<T>(-ph.c1 * delta2 * theta, sig);
            }
            at(C.s.kap[n]) = fmax(T(0), kap);
        }
#undef A4
    }
}

template <class T> bool tiled_closure_supported(const Ctx<T> &C) {
    const DevGrid<T> &g = C.g;
    return (C.ph.closure == OM_CLOSURE_AMD || C.ph.closure == OM_CLOSURE_SMAGORINSKY_LILLY) && g.tx == OM_PERIODIC &&
           g.ty == OM_PERIODIC && g.tz != OM_FLAT && g.Hx >= 1 && g.Hy >= 1 && g.Hz >= 1 && C.ph.n_tracers <= 2 &&
           (!g.kbot || (g.kfix && g.tz == OM_BOUNDED));   // immersed: the generic kernel redoes the cells below kfix
}

template <class T> void launch_closure_tiled(const Ctx<T> &C, int kchunk, hipStream_t st) {
    const DevGrid<T> &g = C.g;
    const int ntx = (g.Nx + CT - 1) / CT, nty = (g.Ny + CT - 1) / CT;
    if (kchunk < 1) kchunk = 1;
    const int nch = (g.Nz + kchunk - 1) / kchunk;
    const dim3 grid(ntx * nty * nch), blk(CT * CT);
    const ClosureTiledArgs<T> A{C, kchunk, ntx};
    switch (C.ph.n_tracers) {
        case 0: launch_k<KClosureTiled<T, 0>>(grid, blk, 0, st, A); break;
        case 1: launch_k<KClosureTiled<T, 1>>(grid, blk, 0, st, A); break;
        default: launch_k<KClosureTiled<T, 2>>(grid, blk, 0, st, A); break;
    }
}

#ifndef OM_ONLY_DOUBLE
template bool tiled_closure_supported<float>(const Ctx<float> &);
template void launch_closure_tiled<float>(const Ctx<float> &, int, hipStream_t);
#endif
#ifndef OM_ONLY_FLOAT
template bool tiled_closure_supported<double>(const Ctx<double> &);
template void launch_closure_tiled<double>(const Ctx<double> &, int, hipStream_t);
#endif

}  // namespace om
