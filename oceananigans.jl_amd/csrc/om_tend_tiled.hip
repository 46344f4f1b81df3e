// om_tend_tiled.hip — z-marching, LDS-tiled tendency kernel for three-dimensional WENO(order=5)
// configurations with Periodic x and y.
//
// Work decomposition.  A block of 256 threads owns an extended tile of 16 × 16 columns and marches
// up a chunk of z levels.  It produces the tendencies of the inner 15 × 15 cells: with that shape
// every x-face flux (16 per row, 15 rows), every y-face flux (15 per column, 16 columns) and every
// top z-face flux of a cell is evaluated exactly once, by one thread, in one pass without
// divergence.  The thread-per-cell form evaluates each face flux twice.
//
// Per z level kk the block stages the planes u(kk), v(kk), tracers(kk), κ(kk), w(kk+1) and ν(kk+1)
// (with halos of 3 in x and y) into LDS; w(kk) and ν(kk−1..kk) are still resident from the previous
// levels.  Each thread keeps its own column's u, v, w and tracers at kk−2 .. kk+3 in registers for
// the z stencils, and carries the top z-face fluxes of level kk to level kk+1 as bottom fluxes.
//
// The flux formulas, their operand order and the tendency assembly are those of k_tendencies
// (om_kernels.hip), which follows nonhydrostatic_tendency_kernel_functions.jl:71-302; the two
// kernels give identically rounded results.
#include <hip/hip_runtime.h>

#include "om_batch.h"
#include "om_kernels.h"

namespace om {

constexpr int TT = 16;            // threads per tile edge
constexpr int TC = 15;            // cells per tile edge
constexpr int PW = TT + 5;        // staged plane width: i0-3 .. i0+17
constexpr int PA = PW * PW;
constexpr int NWD = TT + 1;       // ν / κ plane width: i0-1 .. i0+15
constexpr int NA = NWD * NWD;

template <class T, int NT> struct TileSmem {
    T u[PA], v[PA];
    T c[NT > 0 ? NT : 1][PA];
    T w[2][PA];
    T nu[3][NA];
    T kap[NT > 0 ? NT : 1][NA];
    T fx[2 * (3 + NT)][TT * TT];  // x-face fluxes: [2q] advective, [2q+1] viscous/diffusive
    T fy[2 * (3 + NT)][TT * TT];
};

// z reconstruction at a face from the six values face-3 .. face+2, with the Bounded-z scheme
// chosen for that face (0 WENO5, 1 WENO3, 2 Centered(2); topologically_conditional_interpolation.jl:86-128)
template <class T> __device__ __forceinline__ T zrec(const T (&s)[6], bool left, int sch) {
    if (sch == 0) return weno5<T>(s[0], s[1], s[2], s[3], s[4], s[5], left);
    if (sch == 1) return weno3<T>(s[1], s[2], s[3], s[4], left);
    return cen2<T>(s[2], s[3]);
}

template <class T> __device__ __forceinline__ void stage_plane(T *dst, const T *f, const DevGrid<T> &g, int i0, int j0,
                                                               int k, int tid) {
    const int imax = g.Nx + g.Hx - 1, jmax = g.Ny + g.Hy - 1;
    for (int e = tid; e < PA; e += 256) {
        const int y = e / PW, x = e - y * PW;
        const int i = i0 - 3 + x, j = j0 - 3 + y;
        dst[e] = (i <= imax && j <= jmax) ? f[idx3(i, j, k, g.sy, g.sz)] : T(0);
    }
}

template <class T> __device__ __forceinline__ void stage_small(T *dst, const T *f, const DevGrid<T> &g, int i0, int j0,
                                                               int k, int tid) {
    const int imax = g.Nx + g.Hx - 1, jmax = g.Ny + g.Hy - 1;
    for (int e = tid; e < NA; e += 256) {
        const int y = e / NWD, x = e - y * NWD;
        const int i = i0 - 1 + x, j = j0 - 1 + y;
        dst[e] = (i <= imax && j <= jmax) ? f[idx3(i, j, k, g.sy, g.sz)] : T(0);
    }
}

// Occupancy: with one tracer the FP32 kernel fits 96 VGPRs without spilling, i.e. 5 waves per
// SIMD (5 blocks per CU, 149 KiB of LDS); the hint asks the compiler for that budget.
template <class T, int NT> constexpr int tend_min_blocks() { return (sizeof(T) == 4 && NT <= 1) ? 5 : 1; }

template <class T> struct TendTiledArgs {
    Ctx<T> C;
    int kchunk, ntx;
};

template <class T, int NT> struct KTendTiled {
    static constexpr int kThreads = 256, kMinBlocks = tend_min_blocks<T, NT>();
    static __device__ __forceinline__ void run(const TendTiledArgs<T> &A, const Blk &B);
};

template <class T, int NT>
__device__ __forceinline__ void KTendTiled<T, NT>::run(const TendTiledArgs<T> &A, const Blk &B) {
    __shared__ TileSmem<T, NT> S;
    const Ctx<T> &C = A.C;
    const int kchunk = A.kchunk, ntx = A.ntx;
    const DevGrid<T> &g = C.g;
    const DevPhysics<T> &ph = C.ph;
    const int tid = threadIdx.x;
    const int a = tid & (TT - 1), b = tid >> 4;
    // work item → (z chunk, tile): the chunks of one tile are consecutive, tiles in row order
    const int item = xcd_work_item((int)B.x, (int)B.gx);
    const int nch = (g.Nz + kchunk - 1) / kchunk;
    const int tile = item / nch, chunk = item - tile * nch;
    const int tx = tile % ntx, ty = tile / ntx;
    const int i0 = tx * TC, j0 = ty * TC;
    const int k0 = chunk * kchunk;
    const int k1 = min(k0 + kchunk, g.Nz);
    const int i = i0 + a, j = j0 + b;
    const bool colok = i <= g.Nx + g.Hx - 1 && j <= g.Ny + g.Hy - 1;
    const bool cell = a < TC && b < TC && i < g.Nx && j < g.Ny;
    const int clo = ph.closure;
    const bool has_clo = clo != OM_CLOSURE_NONE;
    const int64_t col = idx3(i, j, 0, g.sy, g.sz);

    // plane accessors: (dx, dy) relative to this thread's column
    auto P = [&](const T *p, int dx, int dy) -> T { return p[(b + 3 + dy) * PW + (a + 3 + dx)]; };
    auto Q = [&](const T *p, int dx, int dy) -> T { return p[(b + 1 + dy) * NWD + (a + 1 + dx)]; };

    // own-column windows: [m] holds level kk-2+m
    T wu[6], wv[6], ww[6], wc[NT > 0 ? NT : 1][6], ko[NT > 0 ? NT : 1][2];
    // own-column loads: a scalar level origin per field plus one per-thread byte offset (the
    // column's offset within a level, < 2³² bytes), shared by every field
    const uint32_t colb0 = (uint32_t)(col * (int64_t)sizeof(T));
    uint32_t colb = colb0;   // refreshed opaquely each level (below), so no per-field pointer is kept live
    auto ldcol = [&](const T *f, int k) -> T {
        return colok ? *reinterpret_cast<const T *>(reinterpret_cast<const char *>(f + (int64_t)k * g.sz) + colb) : T(0);
    };

    int kk = k0 - 1;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        wu[m] = ldcol(C.s.U[0], kk - 2 + m);
        wv[m] = ldcol(C.s.U[1], kk - 2 + m);
        ww[m] = ldcol(C.s.U[2], kk - 2 + m);
#pragma unroll
        for (int n = 0; n < NT; ++n) wc[n][m] = ldcol(C.s.U[3 + n], kk - 2 + m);
    }
    if (clo == OM_CLOSURE_AMD) {
#pragma unroll
        for (int n = 0; n < NT; ++n) { ko[n][0] = ldcol(C.s.kap[n], kk); ko[n][1] = ldcol(C.s.kap[n], kk + 1); }
    }
    // ring slots: w level kk+1 in S.w[wn]; ν levels kk-1, kk, kk+1 in S.nu[(kk+1+r) mod 3]
    int wn = (kk + 1) & 1;
    auto nslot = [&](int level) -> int { return ((level % 3) + 3) % 3; };
    stage_plane(S.w[wn], C.s.U[2], g, i0, j0, kk + 1, tid);
    if (has_clo && clo != OM_CLOSURE_SCALAR_DIFFUSIVITY) {
        stage_small(S.nu[nslot(kk)], C.s.nu, g, i0, j0, kk, tid);
        stage_small(S.nu[nslot(kk + 1)], C.s.nu, g, i0, j0, kk + 1, tid);
    }
    __syncthreads();

    const T Az = g.Az;
    const T half = T(0.5);

    // top z-face fluxes of level `lev` for this column (advective [0], viscous [1]); p13, p23: the
    // products ν·S₁₃ and ν·S₂₃ at this column's fcf and cff points of that z face, which the next
    // level's x- and y-face viscous fluxes of w take over (the same expressions on the same data)
    T zt[3 + NT][2], zb[3 + NT][2];
    T p13 = T(0), p23 = T(0);
    auto ztop = [&](int lev) {
        const T *W1 = S.w[(lev + 1) & 1];
        const T *N0 = S.nu[nslot(lev)], *N1 = S.nu[nslot(lev + 1)];
        const int sf = C.zb_face(lev + 1), sc = C.zb_center(lev);
        // u at fcf face lev+1
        {
            const T wt = acen4<T>(Az, P(W1, -2, 0), P(W1, -1, 0), P(W1, 0, 0), P(W1, 1, 0));
            zt[0][0] = wt * zrec<T>(wu, wt > T(0), sf);
        }
        // v at cff face lev+1
        {
            const T wt = acen4<T>(Az, P(W1, 0, -2), P(W1, 0, -1), P(W1, 0, 0), P(W1, 0, 1));
            zt[1][0] = wt * zrec<T>(wv, wt > T(0), sf);
        }
        // w at ccc centre lev
        {
            const T wt = C.zs4_center(lev) ? acen4<T>(Az, ww[1], ww[2], ww[3], ww[4]) : cen2<T>(Az * ww[2], Az * ww[3]);
            zt[2][0] = wt * zrec<T>(ww, wt > T(0), sc);
        }
#pragma unroll
        for (int n = 0; n < NT; ++n) zt[3 + n][0] = Az * ww[3] * zrec<T>(wc[n], ww[3] > T(0), sf);
        if (has_clo) {
            const T idzf1 = g.idzf[lev + 1];
            T nfcf, ncff, nccc;
            if (clo == OM_CLOSURE_SCALAR_DIFFUSIVITY) { nfcf = ncff = nccc = ph.c0; }
            else {
                nfcf = avg4<T>(Q(N0, -1, 0), Q(N0, 0, 0), Q(N1, -1, 0), Q(N1, 0, 0));
                ncff = avg4<T>(Q(N0, 0, -1), Q(N0, 0, 0), Q(N1, 0, -1), Q(N1, 0, 0));
                nccc = Q(N0, 0, 0);
            }
            const T s13 = half * ((wu[3] - wu[2]) * idzf1 + (P(W1, 0, 0) - P(W1, -1, 0)) * g.idx);
            const T s23 = half * ((wv[3] - wv[2]) * idzf1 + (P(W1, 0, 0) - P(W1, 0, -1)) * g.idy);
            p13 = nfcf * s13;
            p23 = ncff * s23;
            zt[0][1] = Az * (T(-2) * p13);
            zt[1][1] = Az * (T(-2) * p23);
            zt[2][1] = Az * (T(-2) * (nccc * ((ww[3] - ww[2]) * g.idzc[lev])));
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                T kf;
                if (clo == OM_CLOSURE_SCALAR_DIFFUSIVITY) kf = ph.c1;
                else if (clo == OM_CLOSURE_AMD) kf = half * (ko[n][0] + ko[n][1]);
                else kf = (half * (Q(N0, 0, 0) + Q(N1, 0, 0))) / ph.c2;
                zt[3 + n][1] = Az * (-kf * ((wc[n][3] - wc[n][2]) * idzf1));
            }
        }
    };

    ztop(kk);
#pragma unroll
    for (int q = 0; q < 3 + NT; ++q) { zb[q][0] = zt[q][0]; zb[q][1] = zt[q][1]; }

    const bool stage_nu = has_clo && clo != OM_CLOSURE_SCALAR_DIFFUSIVITY;
    // Plane values of the next level are loaded into registers while the current level's fluxes
    // are evaluated, so their memory latency is hidden behind that work.  The plane shares are
    // rebuilt for each load from an opaque copy of tid: a few integer operations, where keeping
    // them live across the flux evaluation would cost registers.
    T ru[2], rv[2], rw[2], rc[NT > 0 ? NT : 1][2], rn[2], rk[NT > 0 ? NT : 1][2];
    auto load_level = [&](int lev) {
        int otid = tid;
        asm volatile("" : "+v"(otid));
        const PlaneShare<T, PW> pbig(g, i0, j0, 3, otid);
        const PlaneShare<T, NWD> psmall(g, i0, j0, 1, otid);
        pbig.load(ru, pbig.origin(C.s.U[0], g, i0, j0, 3, lev));
        pbig.load(rv, pbig.origin(C.s.U[1], g, i0, j0, 3, lev));
#pragma unroll
        for (int n = 0; n < NT; ++n) pbig.load(rc[n], pbig.origin(C.s.U[3 + n], g, i0, j0, 3, lev));
        pbig.load(rw, pbig.origin(C.s.U[2], g, i0, j0, 3, lev + 1));
        if (stage_nu) psmall.load(rn, psmall.origin(C.s.nu, g, i0, j0, 1, lev + 1));
        if (clo == OM_CLOSURE_AMD) {
#pragma unroll
            for (int n = 0; n < NT; ++n) psmall.load(rk[n], psmall.origin(C.s.kap[n], g, i0, j0, 1, lev));
        }
    };
    load_level(k0);
    for (kk = k0; kk < k1; ++kk) {
        colb = colb0;
        asm volatile("" : "+v"(colb));
        int otid = tid;
        asm volatile("" : "+v"(otid));
        const PlaneShare<T, PW> pbig(g, i0, j0, 3, otid);
        const PlaneShare<T, NWD> psmall(g, i0, j0, 1, otid);
        // advance the own-column windows to kk-2 .. kk+3
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            wu[m] = wu[m + 1]; wv[m] = wv[m + 1]; ww[m] = ww[m + 1];
#pragma unroll
            for (int n = 0; n < NT; ++n) wc[n][m] = wc[n][m + 1];
        }
        wu[5] = ldcol(C.s.U[0], kk + 3);
        wv[5] = ldcol(C.s.U[1], kk + 3);
        ww[5] = ldcol(C.s.U[2], kk + 3);
#pragma unroll
        for (int n = 0; n < NT; ++n) wc[n][5] = ldcol(C.s.U[3 + n], kk + 3);
        if (clo == OM_CLOSURE_AMD) {
#pragma unroll
            for (int n = 0; n < NT; ++n) { ko[n][0] = ko[n][1]; ko[n][1] = ldcol(C.s.kap[n], kk + 1); }
        }
        __syncthreads();   // previous level's readers are done with the planes and flux arrays
        pbig.store(S.u, ru, tid);
        pbig.store(S.v, rv, tid);
#pragma unroll
        for (int n = 0; n < NT; ++n) pbig.store(S.c[n], rc[n], tid);
        pbig.store(S.w[(kk + 1) & 1], rw, tid);
        if (stage_nu) psmall.store(S.nu[nslot(kk + 1)], rn, tid);
        if (clo == OM_CLOSURE_AMD) {
#pragma unroll
            for (int n = 0; n < NT; ++n) psmall.store(S.kap[n], rk[n], tid);
        }
        __syncthreads();
        if (kk + 1 < k1) load_level(kk + 1);

        const T *U = S.u, *V = S.v, *W0 = S.w[kk & 1];
        const T *N0 = S.nu[nslot(kk)];
        const T Axc = g.dy * g.dzc[kk], Ayc = g.dx * g.dzc[kk];
        const T Axf = g.dy * g.dzf[kk], Ayf = g.dx * g.dzf[kk];
        const bool s4 = C.zs4_face(kk);

        // νₑ at the ffc edge and the strain S12 there serve both the x-face flux of v and the
        // y-face flux of u: evaluated once
        T nffc_e = T(0), s12_e = T(0);
        if (has_clo) {
            nffc_e = clo == OM_CLOSURE_SCALAR_DIFFUSIVITY ? ph.c0
                   : avg4<T>(Q(N0, -1, -1), Q(N0, 0, -1), Q(N0, -1, 0), Q(N0, 0, 0));
            s12_e = half * ((P(U, 0, 0) - P(U, 0, -1)) * g.idy + (P(V, 0, 0) - P(V, -1, 0)) * g.idx);
        }
        // ---------------- x faces (thread a: u flux at centre i0+a-1, others at face i0+a)
        {
            const T A = Axc;
            {   // F_Uu at ccc (i-1): stencil centred one cell to the left
                const T ut = acen4<T>(A, P(U, -2, 0), P(U, -1, 0), P(U, 0, 0), P(U, 1, 0));
                S.fx[0][tid] = ut * weno5<T>(P(U, -3, 0), P(U, -2, 0), P(U, -1, 0), P(U, 0, 0), P(U, 1, 0), P(U, 2, 0), ut > T(0));
            }
            {   // F_Uv at ffc face i
                const T ut = acen4<T>(A, P(U, 0, -2), P(U, 0, -1), P(U, 0, 0), P(U, 0, 1));
                S.fx[2][tid] = ut * weno5<T>(P(V, -3, 0), P(V, -2, 0), P(V, -1, 0), P(V, 0, 0), P(V, 1, 0), P(V, 2, 0), ut > T(0));
            }
            {   // F_Uw at fcf face i, z-face kk (advecting u from this column's window)
                const T ut = s4 ? cen4<T>(g.dy * g.dzc[kk - 2] * wu[0], g.dy * g.dzc[kk - 1] * wu[1], Axc * wu[2], g.dy * g.dzc[kk + 1] * wu[3])
                                : cen2<T>(g.dy * g.dzc[kk - 1] * wu[1], Axc * wu[2]);
                S.fx[4][tid] = ut * weno5<T>(P(W0, -3, 0), P(W0, -2, 0), P(W0, -1, 0), P(W0, 0, 0), P(W0, 1, 0), P(W0, 2, 0), ut > T(0));
            }
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                const T *Cn = S.c[n];
                const T ui = wu[2];
                S.fx[2 * (3 + n)][tid] = A * ui * weno5<T>(P(Cn, -3, 0), P(Cn, -2, 0), P(Cn, -1, 0), P(Cn, 0, 0), P(Cn, 1, 0), P(Cn, 2, 0), ui > T(0));
            }
            if (has_clo) {
                const T nccc = clo == OM_CLOSURE_SCALAR_DIFFUSIVITY ? ph.c0 : Q(N0, -1, 0);
                S.fx[1][tid] = A * (T(-2) * (nccc * ((P(U, 0, 0) - P(U, -1, 0)) * g.idx)));
                S.fx[3][tid] = A * (T(-2) * (nffc_e * s12_e));
                S.fx[5][tid] = Axf * (T(-2) * p13);   // ν·S₁₃ at (face i, z face kk): the previous level's top face
#pragma unroll
                for (int n = 0; n < NT; ++n) {
                    const T *Cn = S.c[n];
                    T kf;
                    if (clo == OM_CLOSURE_SCALAR_DIFFUSIVITY) kf = ph.c1;
                    else if (clo == OM_CLOSURE_AMD) kf = half * (Q(S.kap[n], -1, 0) + Q(S.kap[n], 0, 0));
                    else kf = (half * (Q(N0, -1, 0) + Q(N0, 0, 0))) / ph.c2;
                    S.fx[2 * (3 + n) + 1][tid] = A * (-kf * ((P(Cn, 0, 0) - P(Cn, -1, 0)) * g.idx));
                }
            }
        }
        // ---------------- y faces (thread b: v flux at centre j0+b-1, others at face j0+b)
        {
            const T A = Ayc;
            {   // F_Vu at ffc face j
                const T vt = acen4<T>(A, P(V, -2, 0), P(V, -1, 0), P(V, 0, 0), P(V, 1, 0));
                S.fy[0][tid] = vt * weno5<T>(P(U, 0, -3), P(U, 0, -2), P(U, 0, -1), P(U, 0, 0), P(U, 0, 1), P(U, 0, 2), vt > T(0));
            }
            {   // F_Vv at ccc (j-1)
                const T vt = acen4<T>(A, P(V, 0, -2), P(V, 0, -1), P(V, 0, 0), P(V, 0, 1));
                S.fy[2][tid] = vt * weno5<T>(P(V, 0, -3), P(V, 0, -2), P(V, 0, -1), P(V, 0, 0), P(V, 0, 1), P(V, 0, 2), vt > T(0));
            }
            {   // F_Vw at cff face j, z-face kk
                const T vt = s4 ? cen4<T>(g.dx * g.dzc[kk - 2] * wv[0], g.dx * g.dzc[kk - 1] * wv[1], Ayc * wv[2], g.dx * g.dzc[kk + 1] * wv[3])
                                : cen2<T>(g.dx * g.dzc[kk - 1] * wv[1], Ayc * wv[2]);
                S.fy[4][tid] = vt * weno5<T>(P(W0, 0, -3), P(W0, 0, -2), P(W0, 0, -1), P(W0, 0, 0), P(W0, 0, 1), P(W0, 0, 2), vt > T(0));
            }
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                const T *Cn = S.c[n];
                const T vj = wv[2];
                S.fy[2 * (3 + n)][tid] = A * vj * weno5<T>(P(Cn, 0, -3), P(Cn, 0, -2), P(Cn, 0, -1), P(Cn, 0, 0), P(Cn, 0, 1), P(Cn, 0, 2), vj > T(0));
            }
            if (has_clo) {
                const T nccc = clo == OM_CLOSURE_SCALAR_DIFFUSIVITY ? ph.c0 : Q(N0, 0, -1);
                S.fy[1][tid] = A * (T(-2) * (nffc_e * s12_e));
                S.fy[3][tid] = A * (T(-2) * (nccc * ((P(V, 0, 0) - P(V, 0, -1)) * g.idy)));
                S.fy[5][tid] = Ayf * (T(-2) * p23);   // ν·S₂₃ at (face j, z face kk): the previous level's top face
#pragma unroll
                for (int n = 0; n < NT; ++n) {
                    const T *Cn = S.c[n];
                    T kf;
                    if (clo == OM_CLOSURE_SCALAR_DIFFUSIVITY) kf = ph.c1;
                    else if (clo == OM_CLOSURE_AMD) kf = half * (Q(S.kap[n], 0, -1) + Q(S.kap[n], 0, 0));
                    else kf = (half * (Q(N0, 0, -1) + Q(N0, 0, 0))) / ph.c2;
                    S.fy[2 * (3 + n) + 1][tid] = A * (-kf * ((P(Cn, 0, 0) - P(Cn, 0, -1)) * g.idy));
                }
            }
        }
        // ---------------- top z faces of level kk (this column)
        ztop(kk);
        __syncthreads();

        if (cell) {
            // outputs and pHY′ through scalar level origins and the column byte offset, as ldcol
            const int64_t lo = (int64_t)kk * g.sz;
            auto at = [&](T *f, int64_t shift) -> T & { return *reinterpret_cast<T *>(reinterpret_cast<char *>(f + lo + shift) + colb); };
            const T iVc = recip<T>(g.Az * g.dzc[kk]), iVf = recip<T>(g.Az * g.dzf[kk]);
            const int sx0 = b * TT + a, sx1 = sx0 + 1, sy1 = sx0 + TT;
            auto diffs = [&](int q, T &sx, T &sy, T &sz, int part) {
                sx = S.fx[2 * q + part][sx1] - S.fx[2 * q + part][sx0];
                sy = S.fy[2 * q + part][sy1] - S.fy[2 * q + part][sx0];
                sz = zt[q][part] - zb[q][part];
            };
            T sx, sy, sz;
            // u
            {
                diffs(0, sx, sy, sz, 0);
                T G = T(0) - iVc * ((sx + sy) + sz);
                if (ph.coriolis) {
                    const T A = Ayc;
                    const T ca = half * (A * P(V, -1, 0) + A * P(V, 0, 0));
                    const T cb = half * (A * P(V, -1, 1) + A * P(V, 0, 1));
                    const T I = half * (ca + cb);
                    G += (half * (ph.f + ph.f)) * I * recip<T>(A);
                }
                if (ph.hydrostatic) G -= (C.s.dxpHY ? at(C.s.dxpHY, 0) : at(C.s.pHY, 0) - at(C.s.pHY, -1)) * g.idx;
                if (has_clo) { diffs(0, sx, sy, sz, 1); G -= iVc * ((sx + sy) + sz); }
                at(C.s.Gn[0], 0) = G;
            }
            // v
            {
                diffs(1, sx, sy, sz, 0);
                T G = T(0) - iVc * ((sx + sy) + sz);
                if (ph.coriolis) {
                    const T A = Axc;
                    const T ca = half * (A * P(U, 0, -1) + A * P(U, 1, -1));
                    const T cb = half * (A * P(U, 0, 0) + A * P(U, 1, 0));
                    const T I = half * (ca + cb);
                    G -= (half * (ph.f + ph.f)) * I * recip<T>(A);
                }
                if (ph.hydrostatic) G -= (C.s.dypHY ? at(C.s.dypHY, 0) : at(C.s.pHY, 0) - at(C.s.pHY, -g.sy)) * g.idy;
                if (has_clo) { diffs(1, sx, sy, sz, 1); G -= iVc * ((sx + sy) + sz); }
                at(C.s.Gn[1], 0) = G;
            }
            // w
            {
                diffs(2, sx, sy, sz, 0);
                T G = T(0) - iVf * ((sx + sy) + sz);
                if (!ph.hydrostatic && ph.buoyancy != OM_BUOYANCY_NONE) {
                    T bm = T(0), b0 = T(0);
                    if (ph.buoyancy == OM_BUOYANCY_TRACER) {
#pragma unroll
                        for (int n = 0; n < NT; ++n) if (n == ph.b_tracer) { bm = wc[n][1]; b0 = wc[n][2]; }
                    } else {
                        T Tm = T(0), T0 = T(0), Sm = T(0), S0 = T(0);
#pragma unroll
                        for (int n = 0; n < NT; ++n) {
                            if (n == ph.T_tracer) { Tm = wc[n][1]; T0 = wc[n][2]; }
                            if (n == ph.S_tracer) { Sm = wc[n][1]; S0 = wc[n][2]; }
                        }
                        bm = ph.g * (ph.alpha * Tm - ph.beta * Sm);
                        b0 = ph.g * (ph.alpha * T0 - ph.beta * S0);
                    }
                    G += half * (bm + b0);
                }
                if (has_clo) { diffs(2, sx, sy, sz, 1); G -= iVf * ((sx + sy) + sz); }
                at(C.s.Gn[2], 0) = G;
            }
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                diffs(3 + n, sx, sy, sz, 0);
                T G = T(0) - iVc * ((sx + sy) + sz);
                if (has_clo) { diffs(3 + n, sx, sy, sz, 1); G -= iVc * ((sx + sy) + sz); }
                at(C.s.Gn[3 + n], 0) = G;
            }
        }
#pragma unroll
        for (int q = 0; q < 3 + NT; ++q) { zb[q][0] = zt[q][0]; zb[q][1] = zt[q][1]; }
    }
}

template <class T> bool tiled_tendencies_supported(const Ctx<T> &C) {
    const DevGrid<T> &g = C.g;
    return C.ph.advection == OM_ADV_WENO5 && g.tx == OM_PERIODIC && g.ty == OM_PERIODIC &&
           (g.tz == OM_BOUNDED || g.tz == OM_PERIODIC) && g.Hx >= 3 && g.Hy >= 3 && g.Hz >= 3 &&
           C.ph.n_tracers <= 2 && g.Nz >= 1 &&
           (!g.kbot || (g.kfix && g.tz == OM_BOUNDED));   // immersed: the generic kernel redoes the cells below kfix
}

template <class T> void launch_tend_tiled(const Ctx<T> &C, int kchunk, hipStream_t st) {
    const DevGrid<T> &g = C.g;
    const int ntx = (g.Nx + TC - 1) / TC, nty = (g.Ny + TC - 1) / TC;
    if (kchunk < 1) kchunk = 1;
    const int nch = (g.Nz + kchunk - 1) / kchunk;
    const dim3 grid(ntx * nty * nch), blk(256);
    const TendTiledArgs<T> A{C, kchunk, ntx};
    switch (C.ph.n_tracers) {
        case 0: launch_k<KTendTiled<T, 0>>(grid, blk, 0, st, A); break;
        case 1: launch_k<KTendTiled<T, 1>>(grid, blk, 0, st, A); break;
        default: launch_k<KTendTiled<T, 2>>(grid, blk, 0, st, A); break;
    }
}

#ifndef OM_ONLY_DOUBLE
template bool tiled_tendencies_supported<float>(const Ctx<float> &);
template void launch_tend_tiled<float>(const Ctx<float> &, int, hipStream_t);
#endif
#ifndef OM_ONLY_FLOAT
template bool tiled_tendencies_supported<double>(const Ctx<double> &);
template void launch_tend_tiled<double>(const Ctx<double> &, int, hipStream_t);
#endif

}  // namespace om
