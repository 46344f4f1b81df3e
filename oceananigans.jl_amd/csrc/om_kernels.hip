// om_kernels.hip — HIP kernels (gfx950) for one NonhydrostaticModel RK3 substage:
// halo fills, eddy viscosity/diffusivity, hydrostatic pressure, the fused tendency kernel,
// the RK3 update, and the pressure correction.
//
// Indices are 0-based device indices (i = 0 is the reference's interior index 1).  Boundary
// switching in z compares the reference index kr = k + 1 against the reference's bounds.
#include <hip/hip_runtime.h>

#include "om_internal.h"
#include "om_batch.h"
#include "om_kernels.h"

namespace om {

// ----------------------------------------------------------------------------------------
// Tendency kernel: one thread per cell computes Gu, Gv, Gw and every tracer tendency.
// nonhydrostatic_tendency_kernel_functions.jl:71-302
// ----------------------------------------------------------------------------------------
// (fixup: the block's columns and level come from the fix-up list, see DevGrid::fix_list)
template <class T> __device__ __forceinline__ void cell_block(const DevGrid<T> &g, int fixup, int &i, int &j, int &k) {
    int bx = blockIdx.x, by = blockIdx.y;
    k = blockIdx.z;
    if (fixup) {
        const int2 e = g.fix_list[blockIdx.x];
        bx = e.x & 0xffff;
        by = e.x >> 16;
        k = e.y;
    }
    i = bx * blockDim.x + threadIdx.x;
    j = by * blockDim.y + threadIdx.y;
}

template <class T> __global__ void __launch_bounds__(256) k_tendencies(Ctx<T> C, int fixup) {
    int i, j, k;
    cell_block(C.g, fixup, i, j, k);
    if (i >= C.g.Nx || j >= C.g.Ny) return;
    if (fixup && k >= C.g.kfix[i + C.g.Nx * j]) return;   // the tiled kernel's value stands
    const bool adv = C.ph.advection != OM_ADV_NONE;
    const bool clo = C.ph.closure != OM_CLOSURE_NONE;
    const int64_t o = C.at(i, j, k);

    // On an immersed grid every difference is a conditional δ, fluxes vanish on the immersed
    // periphery and the Coriolis transports are masked (immersed_boundary_grid.jl conditional
    // operators, coriolis_schemes.jl:56-59); without one every predicate below is false.
    const IB<T> ib(C.g);
    // ---------------- u (fcc)
    {
        T G = T(0);
        if (adv) {
            T sx = C.flx() || ib.template zd<0, 0, 0>(i, j, k, i - 1, j, k) ? T(0) : C.F_Uu(i, j, k) - C.F_Uu(i - 1, j, k);
            T sy = ib.template zd<1, 1, 0>(i, j + 1, k, i, j, k) ? T(0) : C.F_Vu(i, j + 1, k) - C.F_Vu(i, j, k);
            T sz = ib.template zd<1, 0, 1>(i, j, k + 1, i, j, k) ? T(0) : C.F_Wu(i, j, k + 1) - C.F_Wu(i, j, k);
            G -= C.iVc(k) * ((sx + sy) + sz);
        }
        if (C.ph.coriolis) {
            const T *v = C.s.U[1];
            const T A = C.Ayc(k);
            auto q = [&](int a, int b) { return ib.template iper<0, 1, 0>(a, b, k) ? T(0) : A * C.ld(v, a, b, k); };
            T a = C.flx() ? q(i, j) : T(0.5) * (q(i - 1, j) + q(i, j));
            T b = C.flx() ? q(i, j + 1) : T(0.5) * (q(i - 1, j + 1) + q(i, j + 1));
            T I = C.fly() ? a : T(0.5) * (a + b);
            G += (T(0.5) * (C.ph.f + C.ph.f)) * I * recip<T>(A);
        }
        if (C.ph.hydrostatic) {
            const T d = C.s.dxpHY ? C.ld(C.s.dxpHY, i, j, k) : C.ld(C.s.pHY, i, j, k) - C.ld(C.s.pHY, i - 1, j, k);
            G -= (ib.template zd<0, 0, 0>(i, j, k, i - 1, j, k) ? T(0) : d) * C.g.idx;
        }
        if (clo) {
            const T Ax = C.Axc(k), Ay = C.Ayc(k), Az = C.g.Az;
            T sx = C.flx() || ib.template zd<0, 0, 0>(i, j, k, i - 1, j, k) ? T(0) : Ax * C.f_ux(i, j, k) - Ax * C.f_ux(i - 1, j, k);
            T sy = C.fly() || ib.template zd<1, 1, 0>(i, j + 1, k, i, j, k) ? T(0) : Ay * C.f_uy(i, j + 1, k) - Ay * C.f_uy(i, j, k);
            T sz = C.flz() || ib.template zd<1, 0, 1>(i, j, k + 1, i, j, k) ? T(0) : Az * C.f_uz(i, j, k + 1) - Az * C.f_uz(i, j, k);
            G -= C.iVc(k) * ((sx + sy) + sz);
        }
        C.s.Gn[0][o] = G;
    }
    // ---------------- v (cfc)
    {
        T G = T(0);
        if (adv) {
            T sx = ib.template zd<1, 1, 0>(i + 1, j, k, i, j, k) ? T(0) : C.F_Uv(i + 1, j, k) - C.F_Uv(i, j, k);
            T sy = C.fly() || ib.template zd<0, 0, 0>(i, j, k, i, j - 1, k) ? T(0) : C.F_Vv(i, j, k) - C.F_Vv(i, j - 1, k);
            T sz = ib.template zd<0, 1, 1>(i, j, k + 1, i, j, k) ? T(0) : C.F_Wv(i, j, k + 1) - C.F_Wv(i, j, k);
            G -= C.iVc(k) * ((sx + sy) + sz);
        }
        if (C.ph.coriolis) {
            const T *u = C.s.U[0];
            const T A = C.Axc(k);
            auto q = [&](int a, int b) { return ib.template iper<1, 0, 0>(a, b, k) ? T(0) : A * C.ld(u, a, b, k); };
            T a = C.flx() ? q(i, j - 1) : T(0.5) * (q(i, j - 1) + q(i + 1, j - 1));
            T b = C.flx() ? q(i, j) : T(0.5) * (q(i, j) + q(i + 1, j));
            T I = C.fly() ? b : T(0.5) * (a + b);
            G -= (T(0.5) * (C.ph.f + C.ph.f)) * I * recip<T>(A);
        }
        if (C.ph.hydrostatic) {
            const T d = C.s.dypHY ? C.ld(C.s.dypHY, i, j, k) : C.ld(C.s.pHY, i, j, k) - C.ld(C.s.pHY, i, j - 1, k);
            G -= (ib.template zd<0, 0, 0>(i, j, k, i, j - 1, k) ? T(0) : d) * C.g.idy;
        }
        if (clo) {
            const T Ax = C.Axc(k), Ay = C.Ayc(k), Az = C.g.Az;
            T sx = C.flx() || ib.template zd<1, 1, 0>(i + 1, j, k, i, j, k) ? T(0) : Ax * C.f_uy(i + 1, j, k) - Ax * C.f_uy(i, j, k);
            T sy = C.fly() || ib.template zd<0, 0, 0>(i, j, k, i, j - 1, k) ? T(0) : Ay * C.f_vy(i, j, k) - Ay * C.f_vy(i, j - 1, k);
            T sz = C.flz() || ib.template zd<0, 1, 1>(i, j, k + 1, i, j, k) ? T(0) : Az * C.f_vz(i, j, k + 1) - Az * C.f_vz(i, j, k);
            G -= C.iVc(k) * ((sx + sy) + sz);
        }
        C.s.Gn[1][o] = G;
    }
    // ---------------- w (ccf)
    {
        T G = T(0);
        if (adv) {
            T sx = ib.template zd<1, 0, 1>(i + 1, j, k, i, j, k) ? T(0) : C.F_Uw(i + 1, j, k) - C.F_Uw(i, j, k);
            T sy = ib.template zd<0, 1, 1>(i, j + 1, k, i, j, k) ? T(0) : C.F_Vw(i, j + 1, k) - C.F_Vw(i, j, k);
            T sz = C.flz() || ib.template zd<0, 0, 0>(i, j, k, i, j, k - 1) ? T(0) : C.F_Ww(i, j, k) - C.F_Ww(i, j, k - 1);
            G -= C.iVf(k) * ((sx + sy) + sz);
        }
        if (!C.ph.hydrostatic && C.ph.buoyancy != OM_BUOYANCY_NONE && !C.flz())
            G += T(0.5) * (C.buoy(i, j, k - 1) + C.buoy(i, j, k));
        if (clo) {
            const T Ax = C.Axf(k), Ay = C.Ayf(k), Az = C.g.Az;
            T sx = C.flx() || ib.template zd<1, 0, 1>(i + 1, j, k, i, j, k) ? T(0) : Ax * C.f_uz(i + 1, j, k) - Ax * C.f_uz(i, j, k);
            T sy = C.fly() || ib.template zd<0, 1, 1>(i, j + 1, k, i, j, k) ? T(0) : Ay * C.f_vz(i, j + 1, k) - Ay * C.f_vz(i, j, k);
            T sz = C.flz() || ib.template zd<0, 0, 0>(i, j, k, i, j, k - 1) ? T(0) : Az * C.f_wz(i, j, k) - Az * C.f_wz(i, j, k - 1);
            G -= C.iVf(k) * ((sx + sy) + sz);
        }
        C.s.Gn[2][o] = G;
    }
    // ---------------- tracers (ccc)
    // compile-time tracer indices: a runtime index into the by-value pointer tables would
    // place the whole context in per-lane scratch memory
#pragma unroll
    for (int n = 0; n < kMaxTracers; ++n) {
        if (n >= C.ph.n_tracers) break;
        const T *c = C.s.U[3 + n];
        T G = T(0);
        if (adv) {
            T sx = C.flx() || ib.template zd<1, 0, 0>(i + 1, j, k, i, j, k) ? T(0) : C.F_cx(c, i + 1, j, k) - C.F_cx(c, i, j, k);
            T sy = C.fly() || ib.template zd<0, 1, 0>(i, j + 1, k, i, j, k) ? T(0) : C.F_cy(c, i, j + 1, k) - C.F_cy(c, i, j, k);
            T sz = C.flz() || ib.template zd<0, 0, 1>(i, j, k + 1, i, j, k) ? T(0) : C.F_cz(c, i, j, k + 1) - C.F_cz(c, i, j, k);
            G -= C.iVc(k) * ((sx + sy) + sz);
        }
        if (clo) {
            const T Ax = C.Axc(k), Ay = C.Ayc(k), Az = C.g.Az;
            T sx = C.flx() || ib.template zd<1, 0, 0>(i + 1, j, k, i, j, k) ? T(0) : Ax * C.q_x(n, c, i + 1, j, k) - Ax * C.q_x(n, c, i, j, k);
            T sy = C.fly() || ib.template zd<0, 1, 0>(i, j + 1, k, i, j, k) ? T(0) : Ay * C.q_y(n, c, i, j + 1, k) - Ay * C.q_y(n, c, i, j, k);
            T sz = C.flz() || ib.template zd<0, 0, 1>(i, j, k + 1, i, j, k) ? T(0) : Az * C.q_z(n, c, i, j, k + 1) - Az * C.q_z(n, c, i, j, k);
            G -= C.iVc(k) * ((sx + sy) + sz);
        }
        C.s.Gn[3 + n][o] = G;
    }
}

// ----------------------------------------------------------------------------------------
// Eddy viscosity / diffusivity (AMD: anisotropic_minimum_dissipation.jl:161-358;
// Smagorinsky-Lilly: smagorinsky.jl:90-104, lilly_coefficient.jl:129-142)
// ----------------------------------------------------------------------------------------

template <class T> __global__ void __launch_bounds__(256) k_closure(Ctx<T> C, int fixup) {
    int i, j, k;
    cell_block(C.g, fixup, i, j, k);
    if (i >= C.g.Nx || j >= C.g.Ny) return;
    if (fixup && k >= C.g.kfix[i + C.g.Nx * j]) return;   // the tiled kernel's value stands
    const int64_t o = C.at(i, j, k);
    const T dxu = C.dx_u(i, j, k), dyv = C.dy_v(i, j, k), dzw = C.dz_w(i, j, k);

    if (C.ph.closure == OM_CLOSURE_SMAGORINSKY_LILLY) {
        T s12[4] = {C.S12(i, j, k), C.S12(i + 1, j, k), C.S12(i, j + 1, k), C.S12(i + 1, j + 1, k)};
        T s13[4] = {C.S13(i, j, k), C.S13(i + 1, j, k), C.S13(i, j, k + 1), C.S13(i + 1, j, k + 1)};
        T s23[4] = {C.S23(i, j, k), C.S23(i, j + 1, k), C.S23(i, j, k + 1), C.S23(i, j + 1, k + 1)};
        T tr = dxu * dxu + dyv * dyv + dzw * dzw;
        T S2 = tr + T(2) * avg4(s12[0] * s12[0], s12[1] * s12[1], s12[2] * s12[2], s12[3] * s12[3])
                  + T(2) * avg4(s13[0] * s13[0], s13[1] * s13[1], s13[2] * s13[2], s13[3] * s13[3])
                  + T(2) * avg4(s23[0] * s23[0], s23[1] * s23[1], s23[2] * s23[2], s23[3] * s23[3]);
        T D3 = (C.g.dx * C.g.dy) * C.g.dzc[k];
        T Df = cbrt(D3);
        T N2 = T(0.5) * (C.dzb(i, j, k) + C.dzb(i, j, k + 1));
        T c0 = C.ph.c0, cb = C.ph.c1;
        T N2p = fmax(T(0), N2);
        T s2 = T(1) - fmin(T(1), cb * N2p / S2);
        T sig = (S2 == T(0)) ? T(0) : sqrt(s2);
        C.s.nu[o] = (sig * (c0 * c0)) * (Df * Df) * sqrt(T(2) * S2);
        return;
    }

    // ---- AMD: normalised gradients at the edge points around the cell
    const T Dfx = T(2) * C.g.dx, Dfy = T(2) * C.g.dy;
    auto Dfz = [&](int kk) { return T(2) * C.g.dzc[kk]; };
    // ffc points (i,j),(i+1,j),(i,j+1),(i+1,j+1)
    T ndxv[4], ndyu[4];
    {
        const int ii[4] = {i, i + 1, i, i + 1}, jj[4] = {j, j, j + 1, j + 1};
        for (int q = 0; q < 4; ++q) {
            ndxv[q] = Dfx / Dfy * C.dx_v(ii[q], jj[q], k);
            ndyu[q] = Dfy / Dfx * C.dy_u(ii[q], jj[q], k);
        }
    }
    // fcf points (i,k),(i+1,k),(i,k+1),(i+1,k+1); also ∂y_w at those points (see cy_uy below)
    T ndxw[4], ndzu[4], ndyw_x[4];
    {
        const int ii[4] = {i, i + 1, i, i + 1}, kk[4] = {k, k, k + 1, k + 1};
        for (int q = 0; q < 4; ++q) {
            ndxw[q] = Dfx / Dfz(kk[q]) * C.dx_w(ii[q], j, kk[q]);
            ndzu[q] = Dfz(kk[q]) / Dfx * C.dz_u(ii[q], j, kk[q]);
            ndyw_x[q] = Dfy / Dfz(kk[q]) * C.dy_w(ii[q], j, kk[q]);
        }
    }
    // cff points (j,k),(j+1,k),(j,k+1),(j+1,k+1)
    T ndyw[4], ndzv[4];
    {
        const int jj[4] = {j, j + 1, j, j + 1}, kk[4] = {k, k, k + 1, k + 1};
        for (int q = 0; q < 4; ++q) {
            ndyw[q] = Dfy / Dfz(kk[q]) * C.dy_w(i, jj[q], kk[q]);
            ndzv[q] = Dfz(kk[q]) / Dfy * C.dz_v(i, jj[q], kk[q]);
        }
    }
    T nS12[4], nS13[4], nS23[4];
    for (int q = 0; q < 4; ++q) {
        nS12[q] = T(0.5) * (ndyu[q] + ndxv[q]);
        nS13[q] = T(0.5) * (ndzu[q] + ndxw[q]);
        nS23[q] = T(0.5) * (ndzv[q] + ndyw[q]);
    }
#define A4(expr) avg4<T>([&](int q) { return expr; }(0), [&](int q) { return expr; }(1), [&](int q) { return expr; }(2), [&](int q) { return expr; }(3))
    const T q2 = dxu * dxu + dyv * dyv + dzw * dzw
               + A4(ndxv[q] * ndxv[q]) + A4(ndyu[q] * ndyu[q])
               + A4(ndxw[q] * ndxw[q]) + A4(ndzu[q] * ndzu[q])
               + A4(ndyw[q] * ndyw[q]) + A4(ndzv[q] * ndzv[q]);
    const T delta2 = T(3) / ((T(1) / (Dfx * Dfx) + T(1) / (Dfy * Dfy)) + T(1) / (Dfz(k) * Dfz(k)));
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
        nu = ft_div<T>(-C.ph.c0 * delta2 * (r - T(0)), q2);
    }
    C.s.nu[o] = fmax(T(0), nu);

    for (int n = 0; n < C.ph.n_tracers; ++n) {
        const T *c = C.s.U[3 + n];
        const T gx0 = Dfx * C.dxc(c, i, j, k), gx1 = Dfx * C.dxc(c, i + 1, j, k);
        const T gy0 = Dfy * C.dyc(c, i, j, k), gy1 = Dfy * C.dyc(c, i, j + 1, k);
        const T gz0 = Dfz(k) * C.dzc(c, i, j, k), gz1 = Dfz(k + 1) * C.dzc(c, i, j, k + 1);
        const T Ix2 = T(0.5) * (gx0 * gx0 + gx1 * gx1), Iy2 = T(0.5) * (gy0 * gy0 + gy1 * gy1), Iz2 = T(0.5) * (gz0 * gz0 + gz1 * gz1);
        const T sig = (Ix2 + Iy2) + Iz2;
        T kap = T(0);
        if (sig != T(0)) {
            const T Ix = T(0.5) * (gx0 + gx1), Iy = T(0.5) * (gy0 + gy1), Iz = T(0.5) * (gz0 + gz1);
            const T cx_ux = dxu * Ix2 + A4(ndxv[q]) * Ix * Iy + A4(ndxw[q]) * Ix * Iz;
            // the reference interpolates ∂y_w with ℑxzᶜᵃᶜ in this term (anisotropic_minimum_dissipation.jl:343)
            const T cy_uy = A4(ndyu[q]) * Iy * Ix + dyv * Iy2 + A4(ndyw_x[q]) * Iy * Iz;
            const T cz_uz = A4(ndzu[q]) * Iz * Ix + A4(ndzv[q]) * Iz * Iy + dzw * Iz2;
            const T theta = (cx_ux + cy_uy) + cz_uz;
            kap = ft_div<T>(-C.ph.c1 * delta2 * theta, sig);
        }
        C.s.kap[n][o] = fmax(T(0), kap);
    }
#undef A4
}

// ----------------------------------------------------------------------------------------
// Hydrostatic pressure anomaly: downward column integral (update_hydrostatic_pressure.jl:11-21)
// ----------------------------------------------------------------------------------------
template <class T> struct KHydrostatic {
    static constexpr int kThreads = 256, kMinBlocks = 1;
    static __device__ __forceinline__ void run(const CtxArgs<T> &A, const Blk &B);
};
template <class T> __device__ __forceinline__ void KHydrostatic<T>::run(const CtxArgs<T> &A, const Blk &Bk) {
    // The column sum cancels between neighbouring columns when ∂x pHY′ is formed; it is
    // accumulated without FMA contraction, in the reference's order.
#pragma clang fp contract(off)
    const Ctx<T> &C = A.C;
    const int i = Bk.x * blockDim.x + threadIdx.x;
    const int j = Bk.y * blockDim.y + threadIdx.y;
    if (i >= C.g.Nx || j >= C.g.Ny) return;
    const int Nz = C.g.Nz;
    T bk = C.buoy(i, j, Nz);                      // halo value b[Nz+1] (0-based index Nz)
    T bkm = C.buoy(i, j, Nz - 1);
    T p = -(T(0.5) * (bkm + bk)) * C.g.dzf[Nz];
    C.s.pHY[C.at(i, j, Nz - 1)] = p;
    // One thread per column gives one wave per SIMD, so the memory latency is hidden within each
    // thread: the buoyancy is loaded in batches of B levels, double-buffered, the next batch's
    // loads issued before the current batch is integrated (a load placed after a store to pHY′
    // cannot be moved above it, as the arrays might alias).
    constexpr int B = sizeof(T) == 4 ? 32 : 16;
    T ba[B], bb[B];
    auto load = [&](T (&buf)[B], int top) {
#pragma unroll
        for (int q = 0; q < B; ++q) buf[q] = top - q >= 0 ? C.buoy(i, j, top - q) : T(0);
    };
    auto integrate = [&](const T (&buf)[B], int top) {
#pragma unroll
        for (int q = 0; q < B; ++q) {
            if (top - q < 0) break;
            bk = bkm;
            bkm = buf[q];
            p -= (T(0.5) * (bkm + bk)) * C.g.dzf[top - q + 1];
            C.s.pHY[C.at(i, j, top - q)] = p;
        }
    };
    int k = Nz - 2;
    load(ba, k);
    for (; k >= 0; k -= 2 * B) {
        load(bb, k - B);
        integrate(ba, k);
        load(ba, k - 2 * B);
        integrate(bb, k - B);
    }
}

// FP32 pHY′ with the column split into z segments, one wave each, for 64 columns of one row j,
// and the horizontal differences pHY′(i) − pHY′(i−1), pHY′(j) − pHY′(j−1) formed as column sums of
// the buoyancy's x and y differences (when the model holds fields for them): the horizontal
// gradient of pHY′ is a difference of two column sums of size ~Σ b Δz that agree to a few FP32 ulps
// where b varies horizontally by ~1e-5 of itself, and differencing the rounded sums would lose
// most of it; b's neighbour differences are exact in FP32 (Sterbenz) and their sums keep FP32's
// relative precision.
// KHydrostaticReg<T, S> (S segments of at most kHyLevels levels, one wave of 64 columns each,
// Nz ≤ S·kHyLevels; S = 8 up to 128 levels, 16 up to 256): each thread forms its
// segment's terms ½(b(m) + b(m+1))·Δzᶠ(m+1) from the top down with their running sums in registers,
// the segment totals are exchanged through LDS, and p(m) = −(terms of the segments above + the
// running sum at m).  KHydrostaticSeg (any Nz, four segments): the same in two passes, the first
// for the totals, the second recomputing the terms.  Both associate the reference's sequential sum
// in segments, so FP32 results differ from the sequential order by rounding only.
constexpr int kHyLevels = 16;

template <class T, int S> struct KHydrostaticReg {
    static constexpr int kThreads = 64 * S, kMinBlocks = 1;
    static __device__ __forceinline__ void run(const CtxArgs<T> &A, const Blk &B);
};
template <class T, int S> __device__ __forceinline__ void KHydrostaticReg<T, S>::run(const CtxArgs<T> &A, const Blk &B) {
    __shared__ T tot[3][S][64];
    const Ctx<T> &C = A.C;
    // a 1-D grid of (64-column strip, row) items in row-major order, each XCD taking one contiguous
    // run of them: a row's y-neighbour row j − 1 was read by the same XCD just before (its L2)
    const int nbx = (C.g.Nx + 63) / 64;
    const int item = xcd_work_item((int)B.x, (int)B.gx);
    const int i = (item % nbx) * 64 + threadIdx.x;
    const int j = item / nbx;
    const int seg = threadIdx.y, nseg = blockDim.y;
    const int Nz = C.g.Nz;
    const int per = (Nz + nseg - 1) / nseg;
    const int lo = seg * per, hi = min(Nz, lo + per);   // levels lo .. hi-1 of this segment
    const bool col = i < C.g.Nx;
    const bool diff = C.s.dxpHY != nullptr;
    T tp[kHyLevels], tx[kHyLevels], ty[kHyLevels];
    T rp = T(0), rx = T(0), ry = T(0);
    if (col) {
        // terms m = hi-1 down to lo; b(m+1) at m = Nz-1 is the halo value b(Nz)
        T bu = C.buoy(i, j, hi);
        T xu = diff ? C.buoy_diff(i, j, hi, 1, 0) : T(0), yu = diff ? C.buoy_diff(i, j, hi, 0, 1) : T(0);
#pragma unroll
        for (int q = 0; q < kHyLevels; ++q) {
            const int m = hi - 1 - q;
            if (m < lo) break;
            const T bm = C.buoy(i, j, m);
            rp += (T(0.5) * (bm + bu)) * C.g.dzf[m + 1];
            tp[q] = rp;
            bu = bm;
            if (diff) {
                const T xm = C.buoy_diff(i, j, m, 1, 0), ym = C.buoy_diff(i, j, m, 0, 1);
                rx += (T(0.5) * (xm + xu)) * C.g.dzf[m + 1];
                ry += (T(0.5) * (ym + yu)) * C.g.dzf[m + 1];
                tx[q] = rx;
                ty[q] = ry;
                xu = xm;
                yu = ym;
            }
        }
    }
    tot[0][seg][threadIdx.x] = rp;
    tot[1][seg][threadIdx.x] = rx;
    tot[2][seg][threadIdx.x] = ry;
    __syncthreads();
    if (!col) return;
    T ap = T(0), ax = T(0), ay = T(0);
    for (int s2 = seg + 1; s2 < nseg; ++s2) {
        ap += tot[0][s2][threadIdx.x];
        ax += tot[1][s2][threadIdx.x];
        ay += tot[2][s2][threadIdx.x];
    }
#pragma unroll
    for (int q = 0; q < kHyLevels; ++q) {
        const int m = hi - 1 - q;
        if (m < lo) break;
        const int64_t o = C.at(i, j, m);
        C.s.pHY[o] = -(ap + tp[q]);
        if (diff) {
            C.s.dxpHY[o] = -(ax + tx[q]);
            C.s.dypHY[o] = -(ay + ty[q]);
        }
    }
}

template <class T> struct KHydrostaticSeg {
    static constexpr int kThreads = 256, kMinBlocks = 1;
    static __device__ __forceinline__ void run(const CtxArgs<T> &A, const Blk &B);
};
template <class T> __device__ __forceinline__ void KHydrostaticSeg<T>::run(const CtxArgs<T> &A, const Blk &B) {
    __shared__ T tot[3][4][64];
    const Ctx<T> &C = A.C;
    const int i = B.x * 64 + threadIdx.x;
    const int j = B.y;
    const int seg = threadIdx.y;
    const int Nz = C.g.Nz;
    const int per = (Nz + 3) / 4;
    const int lo = seg * per, hi = min(Nz, lo + per);   // levels lo .. hi-1 of this segment
    const bool col = i < C.g.Nx;
    const bool diff = C.s.dxpHY != nullptr;
    // terms m = hi-1 down to lo; b(m+1) at m = Nz-1 is the halo value b(Nz)
    T sp = T(0), sx = T(0), sy = T(0);
    if (col) {
        T bu = C.buoy(i, j, hi), xu = diff ? C.buoy_diff(i, j, hi, 1, 0) : T(0), yu = diff ? C.buoy_diff(i, j, hi, 0, 1) : T(0);
#pragma unroll 4
        for (int m = hi - 1; m >= lo; --m) {
            const T bm = C.buoy(i, j, m);
            sp += (T(0.5) * (bm + bu)) * C.g.dzf[m + 1];
            bu = bm;
            if (diff) {
                const T xm = C.buoy_diff(i, j, m, 1, 0), ym = C.buoy_diff(i, j, m, 0, 1);
                sx += (T(0.5) * (xm + xu)) * C.g.dzf[m + 1];
                sy += (T(0.5) * (ym + yu)) * C.g.dzf[m + 1];
                xu = xm;
                yu = ym;
            }
        }
    }
    tot[0][seg][threadIdx.x] = sp;
    tot[1][seg][threadIdx.x] = sx;
    tot[2][seg][threadIdx.x] = sy;
    __syncthreads();
    if (!col) return;
    T ap = T(0), ax = T(0), ay = T(0);
    for (int s2 = seg + 1; s2 < 4; ++s2) {
        ap += tot[0][s2][threadIdx.x];
        ax += tot[1][s2][threadIdx.x];
        ay += tot[2][s2][threadIdx.x];
    }
    T rp = T(0), rx = T(0), ry = T(0);
    T bu = C.buoy(i, j, hi), xu = diff ? C.buoy_diff(i, j, hi, 1, 0) : T(0), yu = diff ? C.buoy_diff(i, j, hi, 0, 1) : T(0);
#pragma unroll 4
    for (int m = hi - 1; m >= lo; --m) {
        const T bm = C.buoy(i, j, m);
        rp += (T(0.5) * (bm + bu)) * C.g.dzf[m + 1];
        bu = bm;
        C.s.pHY[C.at(i, j, m)] = -(ap + rp);
        if (diff) {
            const T xm = C.buoy_diff(i, j, m, 1, 0), ym = C.buoy_diff(i, j, m, 0, 1);
            rx += (T(0.5) * (xm + xu)) * C.g.dzf[m + 1];
            ry += (T(0.5) * (ym + yu)) * C.g.dzf[m + 1];
            xu = xm;
            yu = ym;
            C.s.dxpHY[C.at(i, j, m)] = -(ax + rx);
            C.s.dypHY[C.at(i, j, m)] = -(ay + ry);
        }
    }
}

// ----------------------------------------------------------------------------------------
// Halo fills (fill_halo_regions*.jl).  One launch per stage for up to kMaxFields fields.
// Stage z: Flux/Value/Gradient one-layer fills of Center fields, impenetrable w faces.
// Stage x, y: periodic copies over the whole parent extent of the other axes.
// ----------------------------------------------------------------------------------------

template <class T> __global__ void k_halo_z(DevGrid<T> g, HaloJob<T> J) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int n = blockIdx.z;
    if (i >= g.Nx || j >= g.Ny || n >= J.nf) return;
    T *c = J.f[n];
    const int Nz = g.Nz;
    const int kb = J.kind_bottom[n], kt = J.kind_top[n];
    if (kb == -2) return;
    if (kb == -1) {  // impenetrable w: boundary faces set to zero
        c[idx3(i, j, 0, g.sy, g.sz)] = T(0);
        c[idx3(i, j, Nz, g.sy, g.sz)] = T(0);
        return;
    }
    {
        const T c1 = c[idx3(i, j, 0, g.sy, g.sz)];
        T h = c1;
        if (kb == OM_BC_VALUE || kb == OM_BC_GRADIENT) {
            const T D = g.dzf[0];
            const T grad = kb == OM_BC_GRADIENT ? J.val_bottom[n] : (c1 - J.val_bottom[n]) / (D / T(2));
            h = c1 + grad * (-D);
        }
        c[idx3(i, j, -1, g.sy, g.sz)] = h;
    }
    {
        const T cN = c[idx3(i, j, Nz - 1, g.sy, g.sz)];
        T h = cN;
        if (kt == OM_BC_VALUE || kt == OM_BC_GRADIENT) {
            const T D = g.dzf[Nz];
            const T grad = kt == OM_BC_GRADIENT ? J.val_top[n] : (J.val_top[n] - cN) / (D / T(2));
            h = cN + grad * D;
        }
        c[idx3(i, j, Nz, g.sy, g.sz)] = h;
    }
}

// periodic fill along axis `ax` for all fields in J; thread grid spans the parent extent of the
// two other axes (including their halos) so corners are copied.
template <class T> __global__ void k_halo_periodic(DevGrid<T> g, HaloJob<T> J, int ax) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;   // first other axis (parent index)
    const int b = blockIdx.y * blockDim.y + threadIdx.y;   // second other axis
    const int n = blockIdx.z;
    if (n >= J.nf) return;
    T *c = J.f[n];
    const int pnx = g.Nx + 2 * g.Hx, pny = g.Ny + 2 * g.Hy, pnz = g.Nz + 2 * g.Hz + 1;
    if (ax == 0) {
        if (a >= pny || b >= pnz) return;
        const int j = a - g.Hy, k = b - g.Hz;
        for (int h = 1; h <= g.Hx; ++h) {
            c[idx3(-h, j, k, g.sy, g.sz)] = c[idx3(g.Nx - h, j, k, g.sy, g.sz)];
            c[idx3(g.Nx - 1 + h, j, k, g.sy, g.sz)] = c[idx3(h - 1, j, k, g.sy, g.sz)];
        }
    } else if (ax == 1) {
        if (a >= pnx || b >= pnz) return;
        const int i = a - g.Hx, k = b - g.Hz;
        for (int h = 1; h <= g.Hy; ++h) {
            c[idx3(i, -h, k, g.sy, g.sz)] = c[idx3(i, g.Ny - h, k, g.sy, g.sz)];
            c[idx3(i, g.Ny - 1 + h, k, g.sy, g.sz)] = c[idx3(i, h - 1, k, g.sy, g.sz)];
        }
    } else {
        if (a >= pnx || b >= pny) return;
        const int i = a - g.Hx, j = b - g.Hy;
        for (int h = 1; h <= g.Hz; ++h) {
            c[idx3(i, j, -h, g.sy, g.sz)] = c[idx3(i, j, g.Nz - h, g.sy, g.sz)];
            c[idx3(i, j, g.Nz - 1 + h, g.sy, g.sz)] = c[idx3(i, j, h - 1, g.sy, g.sz)];
        }
    }
}

// ----------------------------------------------------------------------------------------
// All halo regions of all fields of a job in one launch.  The reference fills z first (over the
// interior x-y extent), then the periodic x halos over the whole parent y-z extent, then the
// periodic y halos over the whole parent x-z extent (boundary_condition_ordering.jl:17-46); so the
// final value of any halo point (i, j, k) is the z-filled value of the interior column
// (wrap(i), wrap(j)) at level k.  Each point is computed from interior values only (never from a
// location this launch writes), so the points are independent.
// ----------------------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T halo_zvalue(const DevGrid<T> &g, const HaloJob<T> &J, int n, const T *c, int64_t col, int k) {
    if (g.tz == OM_BOUNDED) {
        const int kb = J.kind_bottom[n], kt = J.kind_top[n];
        if (kb == -2) return c[col + (int64_t)k * g.sz];
        if (kb == -1) return (k == 0 || k == g.Nz) ? T(0) : c[col + (int64_t)k * g.sz];
        if (k == -1) {
            const T c1 = c[col];
            T h = c1;
            if (kb == OM_BC_VALUE || kb == OM_BC_GRADIENT) {
                const T D = g.dzf[0];
                const T grad = kb == OM_BC_GRADIENT ? J.val_bottom[n] : (c1 - J.val_bottom[n]) / (D / T(2));
                h = c1 + grad * (-D);
            }
            return h;
        }
        if (k == g.Nz) {
            const T cN = c[col + (int64_t)(g.Nz - 1) * g.sz];
            T h = cN;
            if (kt == OM_BC_VALUE || kt == OM_BC_GRADIENT) {
                const T D = g.dzf[g.Nz];
                const T grad = kt == OM_BC_GRADIENT ? J.val_top[n] : (J.val_top[n] - cN) / (D / T(2));
                h = cN + grad * D;
            }
            return h;
        }
        return c[col + (int64_t)k * g.sz];
    }
    if (g.tz == OM_PERIODIC) {
        int kk = k;
        if (k < 0) kk = k + g.Nz;
        else if (k >= g.Nz && k < g.Nz + g.Hz) kk = k - g.Nz;
        return c[col + (int64_t)kk * g.sz];
    }
    return c[col + (int64_t)k * g.sz];
}

template <class T> struct HaloArgs {
    DevGrid<T> g;
    HaloJob<T> J;
    int64_t n_ring, n_z;
};
template <class T> struct KHaloFused {
    static constexpr int kThreads = 256, kMinBlocks = 1;
    static __device__ __forceinline__ void run(const HaloArgs<T> &A, const Blk &B);
};
template <class T> __device__ __forceinline__ void KHaloFused<T>::run(const HaloArgs<T> &A, const Blk &B) {
    const DevGrid<T> &g = A.g;
    const HaloJob<T> &J = A.J;
    const int64_t n_ring = A.n_ring, n_z = A.n_z;
    const int n = B.y;
    T *c = J.f[n];
    const int pnx = g.Nx + 2 * g.Hx, pnz = g.Nz + 2 * g.Hz + 1;
    const int64_t t = B.x * (int64_t)blockDim.x + threadIdx.x;
    auto wrap = [](int v, int N) { return v < 0 ? v + N : (v >= N ? v - N : v); };
    if (t < n_ring) {
        // x-y halo columns over the whole parent z extent: rows j < 0 or j ≥ Ny (full parent width),
        // then the side columns i < 0 or i ≥ Nx of the interior rows
        const int64_t per_k = n_ring / pnz;
        const int kk = (int)(t / per_k);
        int64_t r = t - (int64_t)kk * per_k;
        const int k = kk - g.Hz;
        int i, j;
        const int64_t band = (int64_t)g.Hy * pnx;
        if (r < 2 * band) {
            const int top = r >= band;
            if (top) r -= band;
            const int jj = (int)(r / pnx);
            i = (int)(r - (int64_t)jj * pnx) - g.Hx;
            j = top ? g.Ny + jj : jj - g.Hy;
        } else {
            r -= 2 * band;
            const int w2 = 2 * g.Hx;
            const int jj = (int)(r / w2);
            const int q = (int)(r - (int64_t)jj * w2);
            j = jj;
            i = q < g.Hx ? q - g.Hx : g.Nx + (q - g.Hx);
        }
        const int is = wrap(i, g.Nx), js = wrap(j, g.Ny);
        const int64_t src = idx3(is, js, 0, g.sy, g.sz);
        c[idx3(i, j, k, g.sy, g.sz)] = halo_zvalue(g, J, n, c, src, k);
        return;
    }
    const int64_t u = t - n_ring;
    if (u >= n_z) return;
    // z halo levels of the interior columns
    int k;
    const int64_t nxy = (int64_t)g.Nx * g.Ny;
    const int layer = (int)(u / nxy);
    const int64_t ij = u - (int64_t)layer * nxy;
    if (g.tz == OM_BOUNDED) {
        const int kb = J.kind_bottom[n];
        if (kb == -2) return;
        k = layer == 0 ? (kb == -1 ? 0 : -1) : g.Nz;
    } else {
        k = layer < g.Hz ? -1 - layer : g.Nz + (layer - g.Hz);
    }
    const int j = (int)(ij / g.Nx), i = (int)(ij - (int64_t)j * g.Nx);
    const int64_t col = idx3(i, j, 0, g.sy, g.sz);
    c[col + (int64_t)k * g.sz] = halo_zvalue(g, J, n, c, col, k);
}

// ----------------------------------------------------------------------------------------
// RK3 update of all prognostic fields, with the flux-BC tendency contributions added to Gⁿ
// first (compute_flux_bcs.jl:113-162, runge_kutta_3.jl:196-204, nonhydrostatic_rk3_substep.jl:37-46)
// ----------------------------------------------------------------------------------------

// The same update for NF ≤ 5 fields with every load issued before the first store (a load placed
// after a store to U cannot be moved above it, as the arrays might alias, which would expose one
// memory round trip per field).  Same operations, same order, per field.
template <class T> struct RkPack {
    Ctx<T> C;
    RKArgs<T> a;
};
template <class T, int NF> struct KRk3Batched {
    static constexpr int kThreads = 256, kMinBlocks = 1;
    static __device__ __forceinline__ void run(const RkPack<T> &A, const Blk &B);
};
template <class T, int NF> __device__ __forceinline__ void KRk3Batched<T, NF>::run(const RkPack<T> &A, const Blk &B) {
    const Ctx<T> &C = A.C;
    const RKArgs<T> &a = A.a;
    const int i = B.x * blockDim.x + threadIdx.x;
    const int j = B.y * blockDim.y + threadIdx.y;
    const int k = B.z;
    if (i >= C.g.Nx || j >= C.g.Ny) return;
    const int64_t o = C.at(i, j, k);
    const bool bz = C.bz();
    const T Az = C.g.Az;
    T G[NF], Um[NF], Gp[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        G[f] = C.s.Gn[f][o];
        Um[f] = C.s.U[f][o];
        Gp[f] = a.first ? T(0) : C.s.Gm[f][o];
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        T g = G[f];
        if (bz && f != 2) {
            bool changed = false;
            if (k == 0 && C.ph.bc_bottom[f] == OM_BC_FLUX) { g += C.ph.bc_bottom_value[f] * Az / (Az * C.g.dzc[0]); changed = true; }
            if (k == C.g.Nz - 1 && C.ph.bc_top[f] == OM_BC_FLUX) { g -= C.ph.bc_top_value[f] * Az / (Az * C.g.dzc[C.g.Nz - 1]); changed = true; }
            if (changed) C.s.Gn[f][o] = g;
        }
        if (f == 2 && bz && k == 0 && C.g.Nz > 1) continue;   // exclude_periphery for w
        if (a.first) C.s.U[f][o] = Um[f] + a.dt * a.gamma * g;
        else C.s.U[f][o] = Um[f] + a.dt * (a.gamma * g + a.zeta * Gp[f]);
    }
}

template <class T> struct KRk3 {
    static constexpr int kThreads = 256, kMinBlocks = 1;
    static __device__ __forceinline__ void run(const RkPack<T> &A, const Blk &B);
};
template <class T> __device__ __forceinline__ void KRk3<T>::run(const RkPack<T> &A, const Blk &B) {
    const Ctx<T> &C = A.C;
    const RKArgs<T> &a = A.a;
    const int i = B.x * blockDim.x + threadIdx.x;
    const int j = B.y * blockDim.y + threadIdx.y;
    const int k = B.z;
    if (i >= C.g.Nx || j >= C.g.Ny) return;
    const int64_t o = C.at(i, j, k);
    const bool bz = C.bz();
    const T Az = C.g.Az;
    for (int f = 0; f < a.nfields; ++f) {
        T G = C.s.Gn[f][o];
        if (bz && f != 2) {
            bool changed = false;
            if (k == 0 && C.ph.bc_bottom[f] == OM_BC_FLUX) { G += C.ph.bc_bottom_value[f] * Az / (Az * C.g.dzc[0]); changed = true; }
            if (k == C.g.Nz - 1 && C.ph.bc_top[f] == OM_BC_FLUX) { G -= C.ph.bc_top_value[f] * Az / (Az * C.g.dzc[C.g.Nz - 1]); changed = true; }
            if (changed) C.s.Gn[f][o] = G;
        }
        if (f == 2 && bz && k == 0 && C.g.Nz > 1) continue;   // exclude_periphery for w
        T *U = C.s.U[f];
        if (a.first) U[o] += a.dt * a.gamma * G;
        else U[o] += a.dt * (a.gamma * G + a.zeta * C.s.Gm[f][o]);
    }
}

// ----------------------------------------------------------------------------------------
// Pressure correction U -= ∇(p Δt) (pressure_correction.jl:67-73).  The solver has already stored
// pNHS = (p Δt) / max(ϵ, Δt) (pressure_correction.jl:103), so the kernel only reads the pressure
// and never writes it: no thread overwrites a neighbour's value while it is being read.
// ----------------------------------------------------------------------------------------
// Each thread marches KC levels of one column: p(k − 1) is carried in a register instead of being
// fetched again, and the loads of all KC levels are issued before the first store.
constexpr int kCorrectLevels = 8;

template <class T> struct CorrectArgs {
    Ctx<T> C;
    T dtp;
    int phi;
    int wrap;   // periodic x and y, no immersed bottom: neighbours across the edges by wrapped index
};
template <class T> struct KCorrect {
    static constexpr int kThreads = 256, kMinBlocks = 1;
    static __device__ __forceinline__ void run(const CorrectArgs<T> &A, const Blk &B);
};
template <class T> __device__ __forceinline__ void KCorrect<T>::run(const CorrectArgs<T> &A, const Blk &B) {
    const Ctx<T> &C = A.C;
    const T dtp = A.dtp;
    const int phi = A.phi;
    const int i = B.x * blockDim.x + threadIdx.x;
    const int j = B.y * blockDim.y + threadIdx.y;
    const int k0 = B.z * kCorrectLevels;
    if (i >= C.g.Nx || j >= C.g.Ny) return;
    const int n = min(kCorrectLevels, C.g.Nz - k0);
    const T *p = C.s.pNHS;
    const IB<T> ib(C.g);
    T pc[kCorrectLevels], pw[kCorrectLevels], ps[kCorrectLevels], U0[kCorrectLevels], U1[kCorrectLevels], U2[kCorrectLevels];
    // with wrap, the x, y and z neighbours across the domain edges are read where the halo fill of
    // p would copy them from (the periodic image; at a Bounded bottom the zero-gradient copy of
    // level 0), so p's halos need not be filled first; the values are the same
    const int im = A.wrap && i == 0 ? C.g.Nx - 1 : i - 1;
    const int jm = A.wrap && j == 0 ? C.g.Ny - 1 : j - 1;
    const int kb = A.wrap && k0 == 0 ? (C.g.tz == OM_PERIODIC ? C.g.Nz - 1 : 0) : k0 - 1;
    const T pb = p[C.at(i, j, kb)];   // the level below the chunk (a halo at the bottom wall)
#pragma unroll
    for (int q = 0; q < kCorrectLevels; ++q) {
        if (q >= n) break;
        const int k = k0 + q;
        const int64_t o = C.at(i, j, k);
        pc[q] = p[o];
        pw[q] = p[C.at(im, j, k)];
        ps[q] = p[C.at(i, jm, k)];
        U0[q] = C.s.U[0][o];
        U1[q] = C.s.U[1][o];
        U2[q] = C.s.U[2][o];
    }
#pragma unroll
    for (int q = 0; q < kCorrectLevels; ++q) {
        if (q >= n) break;
        const int k = k0 + q;
        const int64_t o = C.at(i, j, k);
        const T pd = q == 0 ? pb : pc[q - 1];
        // on an immersed grid the gradient is a conditional δ of cell-centred pressure
        const T gx = ib.template zd<0, 0, 0>(i, j, k, i - 1, j, k) ? T(0) : pc[q] - pw[q];
        const T gy = ib.template zd<0, 0, 0>(i, j, k, i, j - 1, k) ? T(0) : pc[q] - ps[q];
        const T gz = ib.template zd<0, 0, 0>(i, j, k, i, j, k - 1) ? T(0) : pc[q] - pd;
        if (phi) {   // ∂ϕ (pressure_correction.jl:82-89)
            if (!C.flx()) C.s.U[0][o] = U0[q] - gx * C.g.idx;
            if (!C.fly()) C.s.U[1][o] = U1[q] - gy * C.g.idy;
            if (!C.flz()) C.s.U[2][o] = U2[q] - gz * C.g.idzf[k];
            continue;
        }
        if (!C.flx()) C.s.U[0][o] = U0[q] - (dtp * gx) * C.g.idx;
        if (!C.fly()) C.s.U[1][o] = U1[q] - (dtp * gy) * C.g.idy;
        if (!C.flz()) C.s.U[2][o] = U2[q] - (dtp * gz) * C.g.idzf[k];
    }
}

// mask_immersed_field! (mask_immersed_field.jl:65-80): zero on the immersed periphery of the
// field's location over its interior indices (k = 0 .. Nz for w on a Bounded z)
template <class T> __global__ void __launch_bounds__(256) k_mask(DevGrid<T> g, T *f, int loc, int nk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.Nx || j >= g.Ny || k >= nk) return;
    const IB<T> ib(g);
    bool z;
    if (loc == 1) z = ib.template iper<1, 0, 0>(i, j, k);
    else if (loc == 2) z = ib.template iper<0, 1, 0>(i, j, k);
    else if (loc == 3) z = ib.template iper<0, 0, 1>(i, j, k);
    else z = ib.template iper<0, 0, 0>(i, j, k);
    if (z) f[idx3(i, j, k, g.sy, g.sz)] = T(0);
}

// ∇·U at cell centres (divergence_operators.jl:16-19)
template <class T> __global__ void __launch_bounds__(256) k_divergence(Ctx<T> C, T *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= C.g.Nx || j >= C.g.Ny) return;
    out[(int64_t)i + (int64_t)C.g.Nx * ((int64_t)j + (int64_t)C.g.Ny * k)] = C.divU(i, j, k, i + 1, j + 1);
}

// ----------------------------------------------------------------------------------------
// Host-side launchers
// ----------------------------------------------------------------------------------------
static inline dim3 cell_grid(int Nx, int Ny, int Nz, dim3 blk) {
    return dim3((Nx + blk.x - 1) / blk.x, (Ny + blk.y - 1) / blk.y, Nz);
}

// the generic kernels' block shape; the fix-up list is built for it
static_assert(kGenericBlockX == 64 && kGenericBlockY == 4, "fix-up list geometry");
template <class T> void launch_tendencies(const Ctx<T> &C, hipStream_t st, bool fixup) {
    dim3 blk(kGenericBlockX, kGenericBlockY, 1);
    if (fixup) {
        if (C.g.fix_count > 0) hipLaunchKernelGGL(k_tendencies<T>, dim3(C.g.fix_count), blk, 0, st, C, 1);
        return;
    }
    hipLaunchKernelGGL(k_tendencies<T>, cell_grid(C.g.Nx, C.g.Ny, C.g.Nz, blk), blk, 0, st, C, 0);
}
template <class T> void launch_closure(const Ctx<T> &C, hipStream_t st, bool fixup) {
    dim3 blk(kGenericBlockX, kGenericBlockY, 1);
    if (fixup) {
        if (C.g.fix_count > 0) hipLaunchKernelGGL(k_closure<T>, dim3(C.g.fix_count), blk, 0, st, C, 1);
        return;
    }
    hipLaunchKernelGGL(k_closure<T>, cell_grid(C.g.Nx, C.g.Ny, C.g.Nz, blk), blk, 0, st, C, 0);
}
template <class T> void launch_hydrostatic(const Ctx<T> &C, hipStream_t st) {
    if (sizeof(T) == 4 && C.g.Nz <= 8 * kHyLevels) {
        // FP32: eight z segments per column scanned in parallel, running sums in registers (the
        // FP64 order is the reference's), with the horizontal differences of pHY′ where held; the
        // block shape does not depend on the grid, so ensemble members of any size up to 128
        // levels share it
        launch_k<KHydrostaticReg<T, 8>>(dim3((C.g.Nx + 63) / 64 * C.g.Ny), dim3(64, 8, 1), 0, st, CtxArgs<T>{C});
        return;
    }
    if (sizeof(T) == 4 && C.g.Nz <= 16 * kHyLevels) {   // sixteen segments, 1024-thread blocks
        launch_k<KHydrostaticReg<T, 16>>(dim3((C.g.Nx + 63) / 64 * C.g.Ny), dim3(64, 16, 1), 0, st, CtxArgs<T>{C});
        return;
    }
    if (sizeof(T) == 4) {
        // FP32, Nz > 8·kHyLevels: four segments, two passes
        dim3 blk(64, 4, 1);
        launch_k<KHydrostaticSeg<T>>(dim3((C.g.Nx + 63) / 64, C.g.Ny), blk, 0, st, CtxArgs<T>{C});
        return;
    }
    dim3 blk(64, 4, 1);
    launch_k<KHydrostatic<T>>(cell_grid(C.g.Nx, C.g.Ny, 1, blk), blk, 0, st, CtxArgs<T>{C});
}
template <class T> void launch_rk3(const Ctx<T> &C, T dt, T gamma, T zeta, int first, int nfields, hipStream_t st) {
    dim3 blk(64, 4, 1);
    const RkPack<T> A{C, RKArgs<T>{dt, gamma, zeta, first, nfields}};
    const dim3 grid = cell_grid(C.g.Nx, C.g.Ny, C.g.Nz, blk);
    switch (nfields) {
        case 3: launch_k<KRk3Batched<T, 3>>(grid, blk, 0, st, A); return;
        case 4: launch_k<KRk3Batched<T, 4>>(grid, blk, 0, st, A); return;
        case 5: launch_k<KRk3Batched<T, 5>>(grid, blk, 0, st, A); return;
        default: launch_k<KRk3<T>>(grid, blk, 0, st, A);
    }
}
template <class T> void launch_correct(const Ctx<T> &C, T dtp, int phi, hipStream_t st, int wrap) {
    dim3 blk(64, 4, 1);
    launch_k<KCorrect<T>>(cell_grid(C.g.Nx, C.g.Ny, (C.g.Nz + kCorrectLevels - 1) / kCorrectLevels, blk), blk, 0, st,
                          CorrectArgs<T>{C, dtp, phi, wrap});
}
template <class T> void launch_mask(const DevGrid<T> &g, T *f, int loc, hipStream_t st) {
    if (!g.kbot) return;
    const int nk = (loc == 3 && g.tz == OM_BOUNDED) ? g.Nz + 1 : g.Nz;
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(k_mask<T>, cell_grid(g.Nx, g.Ny, nk, blk), blk, 0, st, g, f, loc, nk);
}
template <class T> void launch_divergence(const Ctx<T> &C, T *out, hipStream_t st) {
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(k_divergence<T>, cell_grid(C.g.Nx, C.g.Ny, C.g.Nz, blk), blk, 0, st, C, out);
}
template <class T> void launch_halos(const DevGrid<T> &g, const HaloJob<T> &J, hipStream_t st) {
    if (J.nf == 0) return;
    if (g.tx != OM_BOUNDED && g.ty != OM_BOUNDED) {
        const int64_t pnx = g.Nx + 2 * g.Hx, pny = g.Ny + 2 * g.Hy, pnz = g.Nz + 2 * g.Hz + 1;
        const int64_t n_ring = (pnx * pny - (int64_t)g.Nx * g.Ny) * pnz;
        const int64_t layers = g.tz == OM_BOUNDED ? 2 : (g.tz == OM_PERIODIC ? 2 * g.Hz : 0);
        const int64_t n_z = layers * g.Nx * g.Ny;
        const int64_t total = n_ring + n_z;
        if (total == 0) return;
        launch_k<KHaloFused<T>>(dim3((unsigned)((total + 255) / 256), J.nf), dim3(256), 0, st, HaloArgs<T>{g, J, n_ring, n_z});
        return;
    }
    if (g.tz == OM_BOUNDED) {
        dim3 blk(64, 4, 1);
        hipLaunchKernelGGL(k_halo_z<T>, dim3((g.Nx + 63) / 64, (g.Ny + 3) / 4, J.nf), blk, 0, st, g, J);
    }
    const int pnx = g.Nx + 2 * g.Hx, pny = g.Ny + 2 * g.Hy, pnz = g.Nz + 2 * g.Hz + 1;
    dim3 blk(64, 4, 1);
    if (g.tz == OM_PERIODIC && g.Hz > 0)
        hipLaunchKernelGGL(k_halo_periodic<T>, dim3((pnx + 63) / 64, (pny + 3) / 4, J.nf), blk, 0, st, g, J, 2);
    if (g.tx == OM_PERIODIC && g.Hx > 0)
        hipLaunchKernelGGL(k_halo_periodic<T>, dim3((pny + 63) / 64, (pnz + 3) / 4, J.nf), blk, 0, st, g, J, 0);
    if (g.ty == OM_PERIODIC && g.Hy > 0)
        hipLaunchKernelGGL(k_halo_periodic<T>, dim3((pnx + 63) / 64, (pnz + 3) / 4, J.nf), blk, 0, st, g, J, 1);
}

#define OM_INSTANTIATE(T)                                                                           \
    template void launch_tendencies<T>(const Ctx<T> &, hipStream_t, bool);                         \
    template void launch_closure<T>(const Ctx<T> &, hipStream_t, bool);                            \
    template void launch_hydrostatic<T>(const Ctx<T> &, hipStream_t);                              \
    template void launch_rk3<T>(const Ctx<T> &, T, T, T, int, int, hipStream_t);                    \
    template void launch_correct<T>(const Ctx<T> &, T, int, hipStream_t, int);                          \
    template void launch_mask<T>(const DevGrid<T> &, T *, int, hipStream_t);                           \
    template void launch_divergence<T>(const Ctx<T> &, T *, hipStream_t);                          \
    template void launch_halos<T>(const DevGrid<T> &, const HaloJob<T> &, hipStream_t);
// Each precision is compiled in its own object: FP64 without FMA contraction (the reference's
// rounding: Julia never fuses a*b+c implicitly), FP32 with it.
#ifndef OM_ONLY_DOUBLE
OM_INSTANTIATE(float)
#endif
#ifndef OM_ONLY_FLOAT
OM_INSTANTIATE(double)
#endif

}  // namespace om
