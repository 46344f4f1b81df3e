// om_kernels.h — device operators shared by the kernel sources, and the launcher interface
// used by the C-ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>

#include "om_internal.h"

namespace om {

// ----------------------------------------------------------------------------------------
// arithmetic helpers
// ----------------------------------------------------------------------------------------
// Approximate division for the WENO weights (the reference's BackendOptimizedDivision,
// src/Utils/newton_div.jl:57).  FP32: hardware reciprocal (≤1 ulp) times the numerator;
// FP64: IEEE division, so FP64 parity is not limited by it.
template <class T> __device__ __forceinline__ T wdiv(T a, T b);
template <> __device__ __forceinline__ float wdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <> __device__ __forceinline__ double wdiv<double>(double a, double b) { return a / b; }

// Unsigned division by a loop-invariant divisor d ≥ 1 with one multiply-high, a subtract and two
// shifts (Granlund & Montgomery, "Division by invariant integers using multiplication", 1994, §4).
struct FastDiv {
    uint32_t d, mul, s1, s2;
};
inline FastDiv make_fastdiv(uint32_t d) {
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    FastDiv f;
    f.d = d;
    f.mul = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
    f.s1 = l < 1 ? l : 1;
    f.s2 = l > 1 ? l - 1 : 0;
    return f;
}
__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f) {
    const uint32_t t1 = __umulhi(f.mul, n);
    return (t1 + ((n - t1) >> f.s1)) >> f.s2;
}

// Reconstruction coefficients rounded to FT; the last entry of each tuple is 1 - Σ(others)
// evaluated in FT, as the reference builds them (reconstruction_coefficients.jl:74-76).
template <class T> struct RC {
    static constexpr T a0 = T(1.0 / 3), a1 = T(5.0 / 6), a2 = T(1) - (a0 + a1);      // WENO5 stencil 0
    static constexpr T b0 = T(-1.0 / 6), b1 = T(5.0 / 6), b2 = T(1) - (b0 + b1);     // stencil 1
    static constexpr T c0 = T(1.0 / 3), c1 = T(-7.0 / 6), c2 = T(1) - (c0 + c1);     // stencil 2
    static constexpr T d0 = T(0.5), d1 = T(1) - d0;                                   // WENO3 stencil 0
    static constexpr T e0 = T(-0.5), e1 = T(1) - e0;                                  // WENO3 stencil 1
    static constexpr T m1 = T(-1.0 / 12), s7 = T(7.0 / 12);
    static constexpr T q_p1 = m1, q_0 = s7, q_m1 = s7, q_m2 = T(1) - ((m1 + s7) + s7); // Centered(4)
    static constexpr T h_0 = T(0.5), h_m1 = T(1) - T(0.5);                            // Centered(2)
    static constexpr T C50 = T(3.0 / 10), C51 = T(3.0 / 5), C52 = T(1.0 / 10);
    static constexpr T C30 = T(2.0 / 3), C31 = T(1.0 / 3);
    static constexpr T eps = T(1e-8f);
};

// The reference evaluates the smoothness indicators, the WENO weighted sum and the Centered
// interpolants under @muladd (weno_interpolants.jl:265,279,516; centered_reconstruction.jl:56-62):
// a sum of products p₁ + p₂ + … + pₙ, evaluated left to right, becomes
// muladd(pₙ, muladd(…, muladd(p₂, p₁))), i.e. the first product rounded on its own and every later
// one fused into the running sum, and a product a·b·c splits as (a·b, c).  The generic templates
// below (the FP64 path) spell these fused multiply-adds out and contract nothing else, so that
// identical inputs give identically rounded values; FP64 smoothness indicators on raw stencil
// values cancel heavily when the values carry a large offset, which makes the order observable.

// Smoothness quadratic form ψ₁(k₀ψ₁ + k₁ψ₂ + k₂ψ₃) + ψ₂(k₃ψ₂ + k₄ψ₃) + ψ₃ψ₃k₅
// (weno_interpolants.jl:172-174, 228-265)
template <class T> __device__ __forceinline__ T beta_q(T a, T b, T c, T k0, T k1, T k2, T k3, T k4, T k5) {
#pragma clang fp contract(off)
    const T s1 = fma(k2, c, fma(k1, b, k0 * a));
    const T s2 = fma(k4, c, k3 * b);
    return fma(c * c, k5, fma(b, s2, a * s1));
}

// WENO(order=5) value at face i from s0..s5 = ψ[i-3..i+2] (weno_interpolants.jl:348-354, 520-531)
template <class T> __device__ __forceinline__ T weno5(T s0, T s1, T s2, T s3, T s4, T s5, bool left) {
    typedef RC<T> K;
    const T a0 = left ? s2 : s3, a1 = left ? s3 : s2, a2 = left ? s4 : s1;
    const T b0 = left ? s1 : s4, b1 = left ? s2 : s3, b2 = left ? s3 : s2;
    const T c0 = left ? s0 : s5, c1 = left ? s1 : s4, c2 = left ? s2 : s3;
    T ba, bb, bc;
    if (sizeof(T) == 4) {
        // FP32: the central value of each stencil is subtracted first (weno_interpolants.jl:276-281)
        ba = beta_q<T>(a0 - a1, T(0), a2 - a1, T(10), T(-31), T(11), T(25), T(-19), T(4));
        bb = beta_q<T>(b0 - b1, T(0), b2 - b1, T(4), T(-13), T(5), T(13), T(-13), T(4));
        bc = beta_q<T>(c0 - c1, T(0), c2 - c1, T(4), T(-19), T(11), T(25), T(-31), T(10));
    } else {
        ba = beta_q<T>(a0, a1, a2, T(10), T(-31), T(11), T(25), T(-19), T(4));
        bb = beta_q<T>(b0, b1, b2, T(4), T(-13), T(5), T(13), T(-13), T(4));
        bc = beta_q<T>(c0, c1, c2, T(4), T(-19), T(11), T(25), T(-31), T(10));
    }
    const T tau = fabs(ba - bc);
    const T qa = wdiv<T>(tau, ba + K::eps), qb = wdiv<T>(tau, bb + K::eps), qc = wdiv<T>(tau, bc + K::eps);
    const T al0 = K::C50 * (T(1) + qa * qa), al1 = K::C51 * (T(1) + qb * qb), al2 = K::C52 * (T(1) + qc * qc);
    const T sinv = T(1) / ((al0 + al1) + al2);
    // candidate values sum(coeff .* ψ) (no @muladd), then the weighted sum under @muladd
    const T pa = K::a0 * a0 + K::a1 * a1 + K::a2 * a2;
    const T pb = K::b0 * b0 + K::b1 * b1 + K::b2 * b2;
    const T pc = K::c0 * c0 + K::c1 * c1 + K::c2 * c2;
    return fma(al2 * sinv, pc, fma(al1 * sinv, pb, (al0 * sinv) * pa));
}

// FP32 WENO(order=5): the same stencils, smoothness indicators on central-value-shifted values
// and Z weights, evaluated with fused multiply-adds and hardware reciprocals (≤ 1 ulp), and the
// weighted sum normalised once: Σ αₖ pₖ / Σ αₖ.  The reference's FP32 path rounds differently at
// the level of single ulps; FP32 results are held to the FP64 oracle within a stated tolerance,
// never bitwise, so the kernel spends its VALU budget on fewer instructions instead.
template <> __device__ __forceinline__ float weno5<float>(float s0, float s1, float s2, float s3, float s4, float s5, bool left) {
    typedef RC<float> K;
    // Mirror the stencil once: for either bias the three candidate stencils are then
    // a = (t2, t3, t4), b = (t1, t2, t3), c = (t0, t1, t2).
    const float t0 = left ? s0 : s5, t1 = left ? s1 : s4, t2 = left ? s2 : s3, t3 = left ? s3 : s2, t4 = left ? s4 : s1;
    // differences from each stencil's middle value (weno_interpolants.jl:276-281); neighbouring
    // stencils share them: b's right difference is −xa, c's is −xb
    const float xa = t2 - t3, za = t4 - t3, xb = t1 - t2, xc = t0 - t1;
    // The smoothness indicators and ε scaled by 1/4 (exact), which leaves every τ / (β + ε) as it
    // is: β_a/4 = 2.5 xa² + za (2.75 xa + za), β_b/4 = xb (xb − 1.25 xa) + xa²,
    // β_c/4 = 2.5 xb² + xc (xc − 2.75 xb), with xa² and xb² each formed once
    const float sa = xa * xa, sb = xb * xb;
    const float ba = fmaf(2.5f, sa, za * fmaf(2.75f, xa, za));
    const float bb = fmaf(xb, fmaf(-1.25f, xa, xb), sa);
    const float bc = fmaf(2.5f, sb, xc * fmaf(-2.75f, xb, xc));
    const float tau = fabsf(ba - bc);
    constexpr float eps4 = K::eps * 0.25f;
    const float qa = tau * __builtin_amdgcn_rcpf(ba + eps4);
    const float qb = tau * __builtin_amdgcn_rcpf(bb + eps4);
    const float qc = tau * __builtin_amdgcn_rcpf(bc + eps4);
    // the weights are normalised at the end, so the three αₖ may share any factor: here 1 / C₅₁,
    // which saves the middle stencil's multiplication (C₅₀ / C₅₁ = 1/2, C₅₂ / C₅₁ = 1/6)
    const float al0 = 0.5f * fmaf(qa, qa, 1.0f), al1 = fmaf(qb, qb, 1.0f), al2 = (1.0f / 6.0f) * fmaf(qc, qc, 1.0f);
    // candidate values: each stencil's coefficients sum to 1, so p = middle + k₀·x + k₂·z
    const float pa = fmaf(K::a0, xa, fmaf(K::a2, za, t3));
    const float pb = fmaf(K::b0, xb, fmaf(-K::b2, xa, t2));
    const float pc = fmaf(K::c0, xc, fmaf(-K::c2, xb, t1));
    const float num = fmaf(al0, pa, fmaf(al1, pb, al2 * pc));
    return num * __builtin_amdgcn_rcpf((al0 + al1) + al2);
}

// 1 / x: IEEE division in FP64 (the reference's rounding); the hardware reciprocal in FP32
template <class T> __device__ __forceinline__ T recip(T x) { return T(1) / x; }
template <> __device__ __forceinline__ float recip<float>(float x) { return __builtin_amdgcn_rcpf(x); }
// a / b: IEEE in FP64, a times the hardware reciprocal of b in FP32
template <class T> __device__ __forceinline__ T ft_div(T a, T b) { return a / b; }
template <> __device__ __forceinline__ float ft_div<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// WENO(order=3) smoothness ψ₁(ψ₁ − 2ψ₂) + ψ₂ψ₂·1 under @muladd
template <class T> __device__ __forceinline__ T beta2_q(T a, T b) {
#pragma clang fp contract(off)
    return fma(b * b, T(1), a * fma(T(-2), b, T(1) * a));
}

// WENO(order=3) value at face i from s0..s3 = ψ[i-2..i+1]
template <class T> __device__ __forceinline__ T weno3(T s0, T s1, T s2, T s3, bool left) {
    typedef RC<T> K;
    const T a0 = left ? s1 : s2, a1 = left ? s2 : s1;
    const T b0 = left ? s0 : s3, b1 = left ? s1 : s2;
    T ba, bb;
    if (sizeof(T) == 4) {
        const T da = a0 - a1, db = b0 - b1;
        ba = da * (T(1) * da + T(-2) * T(0)) + T(0);
        bb = db * (T(1) * db + T(-2) * T(0)) + T(0);
    } else {
        ba = beta2_q<T>(a0, a1);
        bb = beta2_q<T>(b0, b1);
    }
    const T tau = fabs(ba - bb);
    const T qa = wdiv<T>(tau, ba + K::eps), qb = wdiv<T>(tau, bb + K::eps);
    const T al0 = K::C30 * (T(1) + qa * qa), al1 = K::C31 * (T(1) + qb * qb);
    const T sinv = recip<T>(al0 + al1);
    return fma(al1 * sinv, K::e0 * b0 + K::e1 * b1, (al0 * sinv) * (K::d0 * a0 + K::d1 * a1));
}

// ℑ of four values, as the reference's nested two-point averages
template <class T> __device__ __forceinline__ T avg4(T a, T b, T c, T d) { return T(0.5) * (T(0.5) * (a + b) + T(0.5) * (c + d)); }
// FP32: ¼((a + b) + (c + d)), the same rounded value (scalings by powers of two are exact and
// commute with rounding, subnormal results aside) in two multiplications fewer
template <> __device__ __forceinline__ float avg4<float>(float a, float b, float c, float d) { return 0.25f * ((a + b) + (c + d)); }

// Centered(4) on offsets (−2, −1, 0, +1) from the face, under @muladd
template <class T> __device__ __forceinline__ T cen4(T m2, T m1, T z0, T p1) {
    typedef RC<T> K;
    return fma(K::q_p1, p1, fma(K::q_0, z0, fma(K::q_m1, m1, K::q_m2 * m2)));
}
// cen4 of four values that share one area factor A: in FP64 the reference's order,
// cen4(A·m2, A·m1, A·z0, A·p1); in FP32 the factor applied once to the interpolated value.
template <class T> __device__ __forceinline__ T acen4(T A, T m2, T m1, T z0, T p1) {
    if (sizeof(T) == 4) return A * cen4<T>(m2, m1, z0, p1);
    return cen4<T>(A * m2, A * m1, A * z0, A * p1);
}
// Centered(2) under @muladd (the same value as unfused: both products are exact halvings)
template <class T> __device__ __forceinline__ T cen2(T m1, T z0) {
    typedef RC<T> K;
    return fma(K::h_0, z0, K::h_m1 * m1);
}

// XCD-aware work order for a 1-D grid of nb blocks.  The dispatcher hands consecutive block ids to
// the 8 XCDs in turn, and each XCD has its own L2; this maps the blocks of one XCD onto one
// contiguous run of work items, so that tiles sharing halo planes share an L2.  A bijection on
// [0, nb): the first 8·⌊nb/8⌋ ids are redistributed, the remainder keeps its own id.
__device__ __forceinline__ int xcd_work_item(int id, int nb) {
    const int per = nb >> 3;
    if (id >= (per << 3)) return id;
    return (id & 7) * per + (id >> 3);
}

// A thread's share of a W × W plane staged into LDS by a 256-thread block: plane elements tid and
// tid + 256 (row-major, origin (i0 − h, j0 − h)).  Their byte offsets from the plane origin and
// their validity are computed once per block, so staging a level costs one scalar origin per
// field and a load per element, and every load of a level can be issued before the first LDS
// store.  Elements beyond the field's last halo point in x or y are staged as 0.
template <class T, int W> struct PlaneShare {
    uint32_t off[2];
    bool in1;       // element tid + 256 lies inside the plane (element tid always does: W² ≥ 256)
    bool ok[2];     // inside the plane and the field
    __device__ __forceinline__ PlaneShare(const DevGrid<T> &g, int i0, int j0, int h, int tid) {
        static_assert(W * W >= 256 && W * W <= 512, "two elements per thread");
        const int imax = g.Nx + g.Hx - 1, jmax = g.Ny + g.Hy - 1;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = tid + 256 * u;
            const int y = e / W, x = e - y * W;
            const bool in = e < W * W;
            ok[u] = in && i0 - h + x <= imax && j0 - h + y <= jmax;
            off[u] = (uint32_t)(((int64_t)y * g.sy + x) * (int64_t)sizeof(T));
            if (u == 1) in1 = in;
        }
    }
    // plane origin of field f at level k (uniform across the block)
    __device__ __forceinline__ static const char *origin(const T *f, const DevGrid<T> &g, int i0, int j0, int h, int k) {
        return reinterpret_cast<const char *>(f + idx3(i0 - h, j0 - h, k, g.sy, g.sz));
    }
    __device__ __forceinline__ void load(T (&v)[2], const char *base) const {
#pragma unroll
        for (int u = 0; u < 2; ++u) v[u] = ok[u] ? *reinterpret_cast<const T *>(base + off[u]) : T(0);
    }
    __device__ __forceinline__ void store(T *dst, const T (&v)[2], int tid) const {
        dst[tid] = v[0];
        if (in1) dst[tid + 256] = v[1];
    }
};

// ----------------------------------------------------------------------------------------
// Immersed-bottom predicates (GridFittedBottom; src/Grids/inactive_node.jl:36-137,
// immersed_boundary_interface.jl), device indices (0-based).  A node at location (LX, LY, LZ)
// (0 Center, 1 Face) touches the cells i-a, j-b, k-c for a ≤ LX, b ≤ LY, c ≤ LZ: it is inactive
// when all of them are and peripheral when any of them is (the reference's recursion, unrolled).
// ----------------------------------------------------------------------------------------
template <class T> struct IB {
    const DevGrid<T> &g;
    __device__ __forceinline__ explicit IB(const DevGrid<T> &gg) : g(gg) {}
    __device__ __forceinline__ bool on() const { return g.kbot != nullptr; }
    __device__ __forceinline__ bool outside(int k) const { return g.tz == OM_BOUNDED && (k < 0 || k >= g.Nz); }
    // periodic wrap of a column index into [0, n): one add or subtract for the stencil offsets
    // (|offset| < n), an integer remainder only for points further out (grids narrower than a stencil)
    static __device__ __forceinline__ int wrap(int i, int n) {
        if (i < 0) i += n;
        else if (i >= n) i -= n;
        if ((unsigned)i >= (unsigned)n) {
            i %= n;
            if (i < 0) i += n;
        }
        return i;
    }
    __device__ __forceinline__ bool immersed(int i, int j, int k) const {
        if (!g.kbot || k >= g.kbot_top) return false;   // above the highest bottom: no column lookup
        return k < g.kbot[wrap(i, g.Nx) + g.Nx * wrap(j, g.Ny)];
    }
    __device__ __forceinline__ bool inactive_cell(int i, int j, int k) const { return immersed(i, j, k) || outside(k); }
    template <int LX, int LY, int LZ> __device__ __forceinline__ bool all_inactive(int i, int j, int k, bool under) const {
        bool r = true;
#pragma unroll
        for (int a = 0; a <= LX; ++a)
#pragma unroll
            for (int b = 0; b <= LY; ++b)
#pragma unroll
                for (int c = 0; c <= LZ; ++c) r = r && (under ? outside(k - c) : inactive_cell(i - a, j - b, k - c));
        return r;
    }
    template <int LX, int LY, int LZ> __device__ __forceinline__ bool any_inactive(int i, int j, int k, bool under) const {
        bool r = false;
#pragma unroll
        for (int a = 0; a <= LX; ++a)
#pragma unroll
            for (int b = 0; b <= LY; ++b)
#pragma unroll
                for (int c = 0; c <= LZ; ++c) r = r || (under ? outside(k - c) : inactive_cell(i - a, j - b, k - c));
        return r;
    }
    // immersed_inactive_node / immersed_peripheral_node: on the immersed grid but not on the underlying one
    template <int LX, int LY, int LZ> __device__ __forceinline__ bool iin(int i, int j, int k) const {
        return on() && all_inactive<LX, LY, LZ>(i, j, k, false) && !all_inactive<LX, LY, LZ>(i, j, k, true);
    }
    template <int LX, int LY, int LZ> __device__ __forceinline__ bool iper(int i, int j, int k) const {
        return on() && any_inactive<LX, LY, LZ>(i, j, k, false) && !any_inactive<LX, LY, LZ>(i, j, k, true);
    }
    // conditional δ between two points of one location (conditional_differences.jl:15-44)
    template <int LX, int LY, int LZ> __device__ __forceinline__ bool zd(int i1, int j1, int k1, int i2, int j2, int k2) const {
        return on() && (iin<LX, LY, LZ>(i1, j1, k1) || iin<LX, LY, LZ>(i2, j2, k2));
    }
    // near_{x,y,z}_immersed_boundary for `buffer` (immersed_advective_fluxes.jl:127-215): any inactive
    // node among the 2·buffer stencil points; cell centres for a face interpolation point (cs = false),
    // faces along d for a centre interpolation point (cs = true)
    __device__ __forceinline__ bool near(int d, bool cs, int buffer, int i, int j, int k) const {
        bool r = false;
        for (int n = 1; n <= 2 * buffer; ++n) {
            const int c = cs ? n - buffer : n - buffer - 1;
            const int ii = i + (d == 0 ? c : 0), jj = j + (d == 1 ? c : 0), kk = k + (d == 2 ? c : 0);
            bool in = inactive_cell(ii, jj, kk);
            if (cs) in = in && inactive_cell(ii - (d == 0), jj - (d == 1), kk - (d == 2));
            r = r || in;
        }
        return r;
    }
    // buffer chain WENO5 (0) → WENO3 (1) → Centered(2) (2); `topo` is the choice without an immersed boundary
    __device__ __forceinline__ int bsch(int d, bool cs, int i, int j, int k, int topo) const {
        if (!on()) return topo;
        if (!near(d, cs, 3, i, j, k)) return 0;
        return near(d, cs, 2, i, j, k) ? 2 : 1;
    }
    // advecting velocity: Centered(4) (true) or Centered(2) (false)
    __device__ __forceinline__ bool s4(int d, bool cs, int i, int j, int k, bool topo) const {
        if (!on()) return topo;
        return !near(d, cs, 3, i, j, k);
    }
};

// ----------------------------------------------------------------------------------------
// Context: grid + state + physics with the reference's operators as member functions
// ----------------------------------------------------------------------------------------
template <class T> struct Ctx {
    DevGrid<T> g;
    StateView<T> s;
    DevPhysics<T> ph;

    __device__ __forceinline__ int64_t at(int i, int j, int k) const { return idx3(i, j, k, g.sy, g.sz); }
    __device__ __forceinline__ T ld(const T *f, int i, int j, int k) const { return f[at(i, j, k)]; }
    __device__ __forceinline__ bool flx() const { return g.tx == OM_FLAT; }
    __device__ __forceinline__ bool fly() const { return g.ty == OM_FLAT; }
    __device__ __forceinline__ bool flz() const { return g.tz == OM_FLAT; }
    __device__ __forceinline__ bool bz() const { return g.tz == OM_BOUNDED; }

    // metric operators (area at z-Center / z-Face, volume)
    __device__ __forceinline__ T Axc(int k) const { return g.dy * g.dzc[k]; }
    __device__ __forceinline__ T Axf(int k) const { return g.dy * g.dzf[k]; }
    __device__ __forceinline__ T Ayc(int k) const { return g.dx * g.dzc[k]; }
    __device__ __forceinline__ T Ayf(int k) const { return g.dx * g.dzf[k]; }
    __device__ __forceinline__ T iVc(int k) const { return recip<T>(g.Az * g.dzc[k]); }
    __device__ __forceinline__ T iVf(int k) const { return recip<T>(g.Az * g.dzf[k]); }

    // z boundary switching: 0 = 5th-order (or Centered4), 1 = 3rd-order, 2 = Centered2
    __device__ __forceinline__ int zb_face(int k) const {
        if (!bz()) return 0;
        const int kr = k + 1, N = g.Nz;
        return (kr >= 4 && kr <= N - 2) ? 0 : ((kr >= 3 && kr <= N - 1) ? 1 : 2);
    }
    __device__ __forceinline__ int zb_center(int k) const {
        if (!bz()) return 0;
        const int kr = k + 1, N = g.Nz;
        return (kr >= 3 && kr <= N - 2) ? 0 : ((kr >= 2 && kr <= N - 1) ? 1 : 2);
    }
    __device__ __forceinline__ bool zs4_face(int k) const { const int kr = k + 1; return !bz() || (kr >= 4 && kr <= g.Nz - 2); }
    __device__ __forceinline__ bool zs4_center(int k) const { const int kr = k + 1; return !bz() || (kr >= 3 && kr <= g.Nz - 2); }

    // ---- biased reconstructions of field f at faces with scheme sch (0 WENO5, 1 WENO3, 2 Centered(2));
    // x and y switch only next to an immersed bottom
    __device__ __forceinline__ T bx_s(const T *f, int i, int j, int k, bool left, int sch) const {
        if (sch == 0) return weno5<T>(ld(f, i - 3, j, k), ld(f, i - 2, j, k), ld(f, i - 1, j, k), ld(f, i, j, k), ld(f, i + 1, j, k), ld(f, i + 2, j, k), left);
        if (sch == 1) return weno3<T>(ld(f, i - 2, j, k), ld(f, i - 1, j, k), ld(f, i, j, k), ld(f, i + 1, j, k), left);
        return cen2<T>(ld(f, i - 1, j, k), ld(f, i, j, k));
    }
    __device__ __forceinline__ T by_s(const T *f, int i, int j, int k, bool left, int sch) const {
        if (sch == 0) return weno5<T>(ld(f, i, j - 3, k), ld(f, i, j - 2, k), ld(f, i, j - 1, k), ld(f, i, j, k), ld(f, i, j + 1, k), ld(f, i, j + 2, k), left);
        if (sch == 1) return weno3<T>(ld(f, i, j - 2, k), ld(f, i, j - 1, k), ld(f, i, j, k), ld(f, i, j + 1, k), left);
        return cen2<T>(ld(f, i, j - 1, k), ld(f, i, j, k));
    }
    // z face k with scheme s (stencil centred on face k)
    __device__ __forceinline__ T bz_s(const T *f, int i, int j, int k, bool left, int sch) const {
        if (sch == 0) return weno5<T>(ld(f, i, j, k - 3), ld(f, i, j, k - 2), ld(f, i, j, k - 1), ld(f, i, j, k), ld(f, i, j, k + 1), ld(f, i, j, k + 2), left);
        if (sch == 1) return weno3<T>(ld(f, i, j, k - 2), ld(f, i, j, k - 1), ld(f, i, j, k), ld(f, i, j, k + 1), left);
        return cen2<T>(ld(f, i, j, k - 1), ld(f, i, j, k));
    }

    // ================= momentum advective fluxes (upwind_biased_advective_fluxes.jl:23-93); on an
    // immersed grid they vanish on the immersed periphery (immersed_advective_fluxes.jl:19-70) and the
    // interpolation order drops next to immersed cells (immersed_advective_fluxes.jl:127-215)
    __device__ __forceinline__ T F_Uu(int i, int j, int k) const {  // ccc
        const T *u = s.U[0];
        if (ph.advection == OM_ADV_CENTERED2) { T a = T(0.5) * (ld(u, i, j, k) + ld(u, i + 1, j, k)); return Axc(k) * a * a; }
        const IB<T> ib(g);
        if (ib.template iper<0, 0, 0>(i, j, k)) return T(0);
        const T A = Axc(k);
        const T ut = ib.s4(0, true, i, j, k, true) ? cen4<T>(A * ld(u, i - 1, j, k), A * ld(u, i, j, k), A * ld(u, i + 1, j, k), A * ld(u, i + 2, j, k))
                                                  : cen2<T>(A * ld(u, i, j, k), A * ld(u, i + 1, j, k));
        return ut * bx_s(u, i + 1, j, k, ut > T(0), ib.bsch(0, true, i, j, k, 0));
    }
    __device__ __forceinline__ T F_Vu(int i, int j, int k) const {  // ffc
        if (fly()) return T(0);
        const T *u = s.U[0], *v = s.U[1];
        if (ph.advection == OM_ADV_CENTERED2) {
            T vi = flx() ? ld(v, i, j, k) : T(0.5) * (ld(v, i - 1, j, k) + ld(v, i, j, k));
            return Ayc(k) * vi * (T(0.5) * (ld(u, i, j - 1, k) + ld(u, i, j, k)));
        }
        const IB<T> ib(g);
        if (ib.template iper<1, 1, 0>(i, j, k)) return T(0);
        const T A = Ayc(k);
        const T vt = flx() ? A * ld(v, i, j, k)
                           : (ib.s4(0, false, i, j, k, true) ? cen4<T>(A * ld(v, i - 2, j, k), A * ld(v, i - 1, j, k), A * ld(v, i, j, k), A * ld(v, i + 1, j, k))
                                                             : cen2<T>(A * ld(v, i - 1, j, k), A * ld(v, i, j, k)));
        return vt * by_s(u, i, j, k, vt > T(0), ib.bsch(1, false, i, j, k, 0));
    }
    __device__ __forceinline__ T F_Wu(int i, int j, int k) const {  // fcf
        if (flz()) return T(0);
        const T *u = s.U[0], *w = s.U[2];
        if (ph.advection == OM_ADV_CENTERED2) {
            T wi = flx() ? ld(w, i, j, k) : T(0.5) * (ld(w, i - 1, j, k) + ld(w, i, j, k));
            return g.Az * wi * (T(0.5) * (ld(u, i, j, k - 1) + ld(u, i, j, k)));
        }
        const IB<T> ib(g);
        if (ib.template iper<1, 0, 1>(i, j, k)) return T(0);
        const T A = g.Az;
        const T wt = flx() ? A * ld(w, i, j, k)
                           : (ib.s4(0, false, i, j, k, true) ? cen4<T>(A * ld(w, i - 2, j, k), A * ld(w, i - 1, j, k), A * ld(w, i, j, k), A * ld(w, i + 1, j, k))
                                                             : cen2<T>(A * ld(w, i - 1, j, k), A * ld(w, i, j, k)));
        return wt * bz_s(u, i, j, k, wt > T(0), ib.bsch(2, false, i, j, k, zb_face(k)));
    }
    __device__ __forceinline__ T F_Uv(int i, int j, int k) const {  // ffc
        if (flx()) return T(0);
        const T *u = s.U[0], *v = s.U[1];
        if (ph.advection == OM_ADV_CENTERED2) {
            T uj = fly() ? ld(u, i, j, k) : T(0.5) * (ld(u, i, j - 1, k) + ld(u, i, j, k));
            return Axc(k) * uj * (T(0.5) * (ld(v, i - 1, j, k) + ld(v, i, j, k)));
        }
        const IB<T> ib(g);
        if (ib.template iper<1, 1, 0>(i, j, k)) return T(0);
        const T A = Axc(k);
        const T ut = fly() ? A * ld(u, i, j, k)
                           : (ib.s4(1, false, i, j, k, true) ? cen4<T>(A * ld(u, i, j - 2, k), A * ld(u, i, j - 1, k), A * ld(u, i, j, k), A * ld(u, i, j + 1, k))
                                                             : cen2<T>(A * ld(u, i, j - 1, k), A * ld(u, i, j, k)));
        return ut * bx_s(v, i, j, k, ut > T(0), ib.bsch(0, false, i, j, k, 0));
    }
    __device__ __forceinline__ T F_Vv(int i, int j, int k) const {  // ccc
        const T *v = s.U[1];
        if (ph.advection == OM_ADV_CENTERED2) { T a = T(0.5) * (ld(v, i, j, k) + ld(v, i, j + 1, k)); return Ayc(k) * a * a; }
        const IB<T> ib(g);
        if (ib.template iper<0, 0, 0>(i, j, k)) return T(0);
        const T A = Ayc(k);
        const T vt = ib.s4(1, true, i, j, k, true) ? cen4<T>(A * ld(v, i, j - 1, k), A * ld(v, i, j, k), A * ld(v, i, j + 1, k), A * ld(v, i, j + 2, k))
                                                  : cen2<T>(A * ld(v, i, j, k), A * ld(v, i, j + 1, k));
        return vt * by_s(v, i, j + 1, k, vt > T(0), ib.bsch(1, true, i, j, k, 0));
    }
    __device__ __forceinline__ T F_Wv(int i, int j, int k) const {  // cff
        if (flz()) return T(0);
        const T *v = s.U[1], *w = s.U[2];
        if (ph.advection == OM_ADV_CENTERED2) {
            T wj = fly() ? ld(w, i, j, k) : T(0.5) * (ld(w, i, j - 1, k) + ld(w, i, j, k));
            return g.Az * wj * (T(0.5) * (ld(v, i, j, k - 1) + ld(v, i, j, k)));
        }
        const IB<T> ib(g);
        if (ib.template iper<0, 1, 1>(i, j, k)) return T(0);
        const T A = g.Az;
        const T wt = fly() ? A * ld(w, i, j, k)
                           : (ib.s4(1, false, i, j, k, true) ? cen4<T>(A * ld(w, i, j - 2, k), A * ld(w, i, j - 1, k), A * ld(w, i, j, k), A * ld(w, i, j + 1, k))
                                                             : cen2<T>(A * ld(w, i, j - 1, k), A * ld(w, i, j, k)));
        return wt * bz_s(v, i, j, k, wt > T(0), ib.bsch(2, false, i, j, k, zb_face(k)));
    }
    __device__ __forceinline__ T F_Uw(int i, int j, int k) const {  // fcf
        if (flx()) return T(0);
        const T *u = s.U[0], *w = s.U[2];
        if (ph.advection == OM_ADV_CENTERED2) {
            T uk = flz() ? ld(u, i, j, k) : T(0.5) * (ld(u, i, j, k - 1) + ld(u, i, j, k));
            return Axf(k) * uk * (T(0.5) * (ld(w, i - 1, j, k) + ld(w, i, j, k)));
        }
        const IB<T> ib(g);
        if (ib.template iper<1, 0, 1>(i, j, k)) return T(0);
        T ut;
        if (flz()) ut = Axc(k) * ld(u, i, j, k);
        else if (ib.s4(2, false, i, j, k, zs4_face(k)))
            ut = cen4<T>(Axc(k - 2) * ld(u, i, j, k - 2), Axc(k - 1) * ld(u, i, j, k - 1), Axc(k) * ld(u, i, j, k), Axc(k + 1) * ld(u, i, j, k + 1));
        else ut = cen2<T>(Axc(k - 1) * ld(u, i, j, k - 1), Axc(k) * ld(u, i, j, k));
        return ut * bx_s(w, i, j, k, ut > T(0), ib.bsch(0, false, i, j, k, 0));
    }
    __device__ __forceinline__ T F_Vw(int i, int j, int k) const {  // cff
        if (fly()) return T(0);
        const T *v = s.U[1], *w = s.U[2];
        if (ph.advection == OM_ADV_CENTERED2) {
            T vk = flz() ? ld(v, i, j, k) : T(0.5) * (ld(v, i, j, k - 1) + ld(v, i, j, k));
            return Ayf(k) * vk * (T(0.5) * (ld(w, i, j - 1, k) + ld(w, i, j, k)));
        }
        const IB<T> ib(g);
        if (ib.template iper<0, 1, 1>(i, j, k)) return T(0);
        T vt;
        if (flz()) vt = Ayc(k) * ld(v, i, j, k);
        else if (ib.s4(2, false, i, j, k, zs4_face(k)))
            vt = cen4<T>(Ayc(k - 2) * ld(v, i, j, k - 2), Ayc(k - 1) * ld(v, i, j, k - 1), Ayc(k) * ld(v, i, j, k), Ayc(k + 1) * ld(v, i, j, k + 1));
        else vt = cen2<T>(Ayc(k - 1) * ld(v, i, j, k - 1), Ayc(k) * ld(v, i, j, k));
        return vt * by_s(w, i, j, k, vt > T(0), ib.bsch(1, false, i, j, k, 0));
    }
    __device__ __forceinline__ T F_Ww(int i, int j, int k) const {  // ccc (centre k uses the face k+1 stencils)
        if (flz()) return T(0);
        const T *w = s.U[2];
        if (ph.advection == OM_ADV_CENTERED2) { T a = T(0.5) * (ld(w, i, j, k) + ld(w, i, j, k + 1)); return g.Az * a * a; }
        const IB<T> ib(g);
        if (ib.template iper<0, 0, 0>(i, j, k)) return T(0);
        const T A = g.Az;
        const T wt = ib.s4(2, true, i, j, k, zs4_center(k)) ? cen4<T>(A * ld(w, i, j, k - 1), A * ld(w, i, j, k), A * ld(w, i, j, k + 1), A * ld(w, i, j, k + 2))
                                                           : cen2<T>(A * ld(w, i, j, k), A * ld(w, i, j, k + 1));
        return wt * bz_s(w, i, j, k + 1, wt > T(0), ib.bsch(2, true, i, j, k, zb_center(k)));
    }
    // ================= tracer advective fluxes (upwind_biased_advective_fluxes.jl:99-121)
    __device__ __forceinline__ T F_cx(const T *c, int i, int j, int k) const {
        const T ui = ld(s.U[0], i, j, k);
        if (ph.advection == OM_ADV_CENTERED2) return (Axc(k) * ui) * cen2<T>(ld(c, i - 1, j, k), ld(c, i, j, k));
        const IB<T> ib(g);
        if (ib.template iper<1, 0, 0>(i, j, k)) return T(0);
        return Axc(k) * ui * bx_s(c, i, j, k, ui > T(0), ib.bsch(0, false, i, j, k, 0));
    }
    __device__ __forceinline__ T F_cy(const T *c, int i, int j, int k) const {
        const T vj = ld(s.U[1], i, j, k);
        if (ph.advection == OM_ADV_CENTERED2) return (Ayc(k) * vj) * cen2<T>(ld(c, i, j - 1, k), ld(c, i, j, k));
        const IB<T> ib(g);
        if (ib.template iper<0, 1, 0>(i, j, k)) return T(0);
        return Ayc(k) * vj * by_s(c, i, j, k, vj > T(0), ib.bsch(1, false, i, j, k, 0));
    }
    __device__ __forceinline__ T F_cz(const T *c, int i, int j, int k) const {
        const T wk = ld(s.U[2], i, j, k);
        if (ph.advection == OM_ADV_CENTERED2) return (g.Az * wk) * cen2<T>(ld(c, i, j, k - 1), ld(c, i, j, k));
        const IB<T> ib(g);
        if (ib.template iper<0, 0, 1>(i, j, k)) return T(0);
        return g.Az * wk * bz_s(c, i, j, k, wk > T(0), ib.bsch(2, false, i, j, k, zb_face(k)));
    }

    // ================= gradients (velocity_tracer_gradients.jl:6-18)
    // conditional δ next to an immersed bottom (conditional_differences.jl:15-44)
    __device__ __forceinline__ T dx_u(int i, int j, int k) const { return flx() || IB<T>(g).template zd<1, 0, 0>(i, j, k, i + 1, j, k) ? T(0) : (ld(s.U[0], i + 1, j, k) - ld(s.U[0], i, j, k)) * g.idx; }
    __device__ __forceinline__ T dy_v(int i, int j, int k) const { return fly() || IB<T>(g).template zd<0, 1, 0>(i, j, k, i, j + 1, k) ? T(0) : (ld(s.U[1], i, j + 1, k) - ld(s.U[1], i, j, k)) * g.idy; }
    __device__ __forceinline__ T dz_w(int i, int j, int k) const { return flz() || IB<T>(g).template zd<0, 0, 1>(i, j, k, i, j, k + 1) ? T(0) : (ld(s.U[2], i, j, k + 1) - ld(s.U[2], i, j, k)) * g.idzc[k]; }
    __device__ __forceinline__ T dx_v(int i, int j, int k) const { return flx() || IB<T>(g).template zd<0, 1, 0>(i, j, k, i - 1, j, k) ? T(0) : (ld(s.U[1], i, j, k) - ld(s.U[1], i - 1, j, k)) * g.idx; }
    __device__ __forceinline__ T dx_w(int i, int j, int k) const { return flx() || IB<T>(g).template zd<0, 0, 1>(i, j, k, i - 1, j, k) ? T(0) : (ld(s.U[2], i, j, k) - ld(s.U[2], i - 1, j, k)) * g.idx; }
    __device__ __forceinline__ T dy_u(int i, int j, int k) const { return fly() || IB<T>(g).template zd<1, 0, 0>(i, j, k, i, j - 1, k) ? T(0) : (ld(s.U[0], i, j, k) - ld(s.U[0], i, j - 1, k)) * g.idy; }
    __device__ __forceinline__ T dy_w(int i, int j, int k) const { return fly() || IB<T>(g).template zd<0, 0, 1>(i, j, k, i, j - 1, k) ? T(0) : (ld(s.U[2], i, j, k) - ld(s.U[2], i, j - 1, k)) * g.idy; }
    __device__ __forceinline__ T dz_u(int i, int j, int k) const { return flz() || IB<T>(g).template zd<1, 0, 0>(i, j, k, i, j, k - 1) ? T(0) : (ld(s.U[0], i, j, k) - ld(s.U[0], i, j, k - 1)) * g.idzf[k]; }
    __device__ __forceinline__ T dz_v(int i, int j, int k) const { return flz() || IB<T>(g).template zd<0, 1, 0>(i, j, k, i, j, k - 1) ? T(0) : (ld(s.U[1], i, j, k) - ld(s.U[1], i, j, k - 1)) * g.idzf[k]; }
    __device__ __forceinline__ T dxc(const T *c, int i, int j, int k) const { return flx() || IB<T>(g).template zd<0, 0, 0>(i, j, k, i - 1, j, k) ? T(0) : (ld(c, i, j, k) - ld(c, i - 1, j, k)) * g.idx; }
    __device__ __forceinline__ T dyc(const T *c, int i, int j, int k) const { return fly() || IB<T>(g).template zd<0, 0, 0>(i, j, k, i, j - 1, k) ? T(0) : (ld(c, i, j, k) - ld(c, i, j - 1, k)) * g.idy; }
    __device__ __forceinline__ T dzc(const T *c, int i, int j, int k) const { return flz() || IB<T>(g).template zd<0, 0, 0>(i, j, k, i, j, k - 1) ? T(0) : (ld(c, i, j, k) - ld(c, i, j, k - 1)) * g.idzf[k]; }
    __device__ __forceinline__ T S12(int i, int j, int k) const { return T(0.5) * (dy_u(i, j, k) + dx_v(i, j, k)); }
    __device__ __forceinline__ T S13(int i, int j, int k) const { return T(0.5) * (dz_u(i, j, k) + dx_w(i, j, k)); }
    __device__ __forceinline__ T S23(int i, int j, int k) const { return T(0.5) * (dz_v(i, j, k) + dy_w(i, j, k)); }

    // ================= eddy viscosity at the stress locations (abstract_scalar_diffusivity_closure.jl:343-353)
    __device__ __forceinline__ T nu_ccc(int i, int j, int k) const {
        if (ph.closure == OM_CLOSURE_SCALAR_DIFFUSIVITY) return ph.c0;
        return ld(s.nu, i, j, k);
    }
    __device__ __forceinline__ T nu_ffc(int i, int j, int k) const {
        if (ph.closure == OM_CLOSURE_SCALAR_DIFFUSIVITY) return ph.c0;
        const T *n = s.nu;
        T a = flx() ? ld(n, i, j - 1, k) : T(0.5) * (ld(n, i - 1, j - 1, k) + ld(n, i, j - 1, k));
        T b = flx() ? ld(n, i, j, k) : T(0.5) * (ld(n, i - 1, j, k) + ld(n, i, j, k));
        if (fly()) return b;
        return T(0.5) * (a + b);
    }
    __device__ __forceinline__ T nu_fcf(int i, int j, int k) const {
        if (ph.closure == OM_CLOSURE_SCALAR_DIFFUSIVITY) return ph.c0;
        const T *n = s.nu;
        T a = flx() ? ld(n, i, j, k - 1) : T(0.5) * (ld(n, i - 1, j, k - 1) + ld(n, i, j, k - 1));
        T b = flx() ? ld(n, i, j, k) : T(0.5) * (ld(n, i - 1, j, k) + ld(n, i, j, k));
        if (flz()) return b;
        return T(0.5) * (a + b);
    }
    __device__ __forceinline__ T nu_cff(int i, int j, int k) const {
        if (ph.closure == OM_CLOSURE_SCALAR_DIFFUSIVITY) return ph.c0;
        const T *n = s.nu;
        T a = fly() ? ld(n, i, j, k - 1) : T(0.5) * (ld(n, i, j - 1, k - 1) + ld(n, i, j, k - 1));
        T b = fly() ? ld(n, i, j, k) : T(0.5) * (ld(n, i, j - 1, k) + ld(n, i, j, k));
        if (flz()) return b;
        return T(0.5) * (a + b);
    }
    // tracer diffusivity at fcc / cfc / ccf
    __device__ __forceinline__ T kap_f(int n, int i, int j, int k, int di, int dj, int dk) const {
        if (ph.closure == OM_CLOSURE_SCALAR_DIFFUSIVITY) return ph.c1;
        if (ph.closure == OM_CLOSURE_AMD) return T(0.5) * (ld(s.kap[n], i - di, j - dj, k - dk) + ld(s.kap[n], i, j, k));
        return (T(0.5) * (ld(s.nu, i - di, j - dj, k - dk) + ld(s.nu, i, j, k))) / ph.c2;
    }

    // ∇·U at a cell centre (divergence_operators.jl:16-19), conditional δ on an immersed grid;
    // ip, jp address the i + 1, j + 1 neighbours (wrapped by a caller whose halos are stale)
    __device__ __forceinline__ T divU(int i, int j, int k, int ip, int jp) const {
        const IB<T> ib(g);
        const T *u = s.U[0], *v = s.U[1], *w = s.U[2];
        T sx = flx() || ib.template zd<1, 0, 0>(i, j, k, i + 1, j, k) ? T(0) : Axc(k) * ld(u, ip, j, k) - Axc(k) * ld(u, i, j, k);
        T sy = fly() || ib.template zd<0, 1, 0>(i, j, k, i, j + 1, k) ? T(0) : Ayc(k) * ld(v, i, jp, k) - Ayc(k) * ld(v, i, j, k);
        T sz = flz() || ib.template zd<0, 0, 1>(i, j, k, i, j, k + 1) ? T(0) : g.Az * ld(w, i, j, k + 1) - g.Az * ld(w, i, j, k);
        return iVc(k) * ((sx + sy) + sz);
    }

    // viscous and diffusive fluxes (abstract_scalar_diffusivity_closure.jl:214-262), zero on the
    // immersed periphery of their location (immersed_diffusive_fluxes.jl:25-42)
    __device__ __forceinline__ T f_ux(int i, int j, int k) const { return IB<T>(g).template iper<0, 0, 0>(i, j, k) ? T(0) : T(-2) * (nu_ccc(i, j, k) * dx_u(i, j, k)); }
    __device__ __forceinline__ T f_uy(int i, int j, int k) const { return IB<T>(g).template iper<1, 1, 0>(i, j, k) ? T(0) : T(-2) * (nu_ffc(i, j, k) * S12(i, j, k)); }
    __device__ __forceinline__ T f_uz(int i, int j, int k) const { return IB<T>(g).template iper<1, 0, 1>(i, j, k) ? T(0) : T(-2) * (nu_fcf(i, j, k) * S13(i, j, k)); }
    __device__ __forceinline__ T f_vy(int i, int j, int k) const { return IB<T>(g).template iper<0, 0, 0>(i, j, k) ? T(0) : T(-2) * (nu_ccc(i, j, k) * dy_v(i, j, k)); }
    __device__ __forceinline__ T f_vz(int i, int j, int k) const { return IB<T>(g).template iper<0, 1, 1>(i, j, k) ? T(0) : T(-2) * (nu_cff(i, j, k) * S23(i, j, k)); }
    __device__ __forceinline__ T f_wz(int i, int j, int k) const { return IB<T>(g).template iper<0, 0, 0>(i, j, k) ? T(0) : T(-2) * (nu_ccc(i, j, k) * dz_w(i, j, k)); }
    __device__ __forceinline__ T q_x(int n, const T *c, int i, int j, int k) const { return IB<T>(g).template iper<1, 0, 0>(i, j, k) ? T(0) : -kap_f(n, i, j, k, 1, 0, 0) * dxc(c, i, j, k); }
    __device__ __forceinline__ T q_y(int n, const T *c, int i, int j, int k) const { return IB<T>(g).template iper<0, 1, 0>(i, j, k) ? T(0) : -kap_f(n, i, j, k, 0, 1, 0) * dyc(c, i, j, k); }
    __device__ __forceinline__ T q_z(int n, const T *c, int i, int j, int k) const { return IB<T>(g).template iper<0, 0, 1>(i, j, k) ? T(0) : -kap_f(n, i, j, k, 0, 0, 1) * dzc(c, i, j, k); }

    __device__ __forceinline__ T buoy(int i, int j, int k) const {
        if (ph.buoyancy == OM_BUOYANCY_TRACER) return ld(s.btr, i, j, k);
        if (ph.buoyancy == OM_BUOYANCY_LINEAR_EOS)
            return ph.g * (ph.alpha * ld(s.Ttr, i, j, k) - ph.beta * ld(s.Str, i, j, k));
        return T(0);
    }
    // b(i, j, k) − b(i − di, j − dj, k), formed from the tracers' differences (exact in FP32 for
    // neighbouring values within a factor two of each other, Sterbenz), not from two rounded b
    __device__ __forceinline__ T buoy_diff(int i, int j, int k, int di, int dj) const {
        if (ph.buoyancy == OM_BUOYANCY_TRACER) return ld(s.btr, i, j, k) - ld(s.btr, i - di, j - dj, k);
        if (ph.buoyancy == OM_BUOYANCY_LINEAR_EOS)
            return ph.g * (ph.alpha * (ld(s.Ttr, i, j, k) - ld(s.Ttr, i - di, j - dj, k)) -
                           ph.beta * (ld(s.Str, i, j, k) - ld(s.Str, i - di, j - dj, k)));
        return T(0);
    }
    __device__ __forceinline__ T dzb(int i, int j, int k) const {  // ∂z_b at ccf
        if (ph.buoyancy == OM_BUOYANCY_TRACER) return dzc(s.btr, i, j, k);
        if (ph.buoyancy == OM_BUOYANCY_LINEAR_EOS)
            return ph.g * (ph.alpha * dzc(s.Ttr, i, j, k) - ph.beta * dzc(s.Str, i, j, k));
        return T(0);
    }
};

template <class T> struct HaloJob {
    T *f[kMaxFields + 4];
    int kind_bottom[kMaxFields + 4], kind_top[kMaxFields + 4];   // OM_BC_* ; -1 = w normal flow; -2 = skip
    T val_bottom[kMaxFields + 4], val_top[kMaxFields + 4];
    int nf;
};

// the argument pack of a kernel that takes only the model context
template <class T> struct CtxArgs {
    Ctx<T> C;
};

template <class T> struct RKArgs {
    T dt, gamma, zeta;
    int first;
    int nfields;
};

// Poisson solver plan: spectral storage and per-axis tables (device pointers).
template <class T> struct PoissonPlan {
    typename Vec2<T>::type *S;          // complex spectral array, Ms · Ny · Nz (x rows padded to Ms)
    int Ms;                              // row stride of S: Mx = Nx/2 + 1 rounded up to a 128-byte line
    const typename Vec2<T>::type *tw_xh; // twiddles for the x line (length Nx/2 when Nx is even, Nx otherwise)
    const typename Vec2<T>::type *post;  // exp(-2πi k / Nx), k = 0 .. Nx/2
    const typename Vec2<T>::type *tw_y;  // length Ny
    const typename Vec2<T>::type *tw_z;  // length Lz_line
    const typename Vec2<T>::type *tw_dct; // exp(-iπk / 2Nz), k < Nz (zmode 1)
    const T *lx, *ly, *lz;               // eigenvalues (lz: Nz entries, 2·Nz for Bounded z)
    int zmode;                           // 0 Periodic z, 1 Bounded z via a length-Nz DCT, 2 Bounded z, odd Nz, 2Nz extension
    int Lz_line;                         // FFT length along z: Nz (zmodes 0, 1), 2·Nz (zmode 2)
    int Lz_norm;                         // normalisation of the z transform pair: Nz (Periodic), 2·Nz (Bounded)
    int tridiagonal;                     // Fourier–tridiagonal z solve (stretched z)
    T eps10;                             // 10·eps(FT): the Thomas pivot guard
};

// Lines of the FFT passes staged per block: the largest power of two R ≤ 16 with R·n ≤ 4096 whose
// R rows (stride n + 1 when R > 1) and `tables` complex table entries fit in 64 KiB of LDS, with a
// second buffer when n is not a power of two (direct DFT).  0: not even one line fits.
inline int fft_lines_per_block(int n, int tables, size_t elem) {
    const bool p2 = (n & (n - 1)) == 0;
    for (int R = 16; R >= 1; R >>= 1) {
        if ((size_t)R * n > 4096) continue;
        const size_t ld = R > 1 ? (size_t)n + 1 : (size_t)n;
        if (((size_t)(p2 ? 1 : 2) * R * ld + (size_t)tables) * elem <= 65536) return R;
    }
    return 0;
}

// The generic kernels over every cell, or (fixup) only over the cells below kfix of an immersed
// grid, after the tiled kernels have covered the rest
constexpr int kImmersedReach = 5;   // farthest offset of an immersed-boundary query from its cell, plus one
constexpr int kGenericBlockX = 64, kGenericBlockY = 4;   // block shape of k_tendencies / k_closure
template <class T> void launch_tendencies(const Ctx<T> &C, hipStream_t st, bool fixup = false);
// z-marching LDS-tiled tendency kernel (om_tend_tiled.hip) and the configurations it covers
template <class T> bool tiled_tendencies_supported(const Ctx<T> &C);
template <class T> void launch_tend_tiled(const Ctx<T> &C, int kchunk, hipStream_t st);
template <class T> void launch_closure(const Ctx<T> &C, hipStream_t st, bool fixup = false);
// z-marching LDS-tiled eddy viscosity / diffusivity kernel (om_closure_tiled.hip)
template <class T> bool tiled_closure_supported(const Ctx<T> &C);
template <class T> void launch_closure_tiled(const Ctx<T> &C, int kchunk, hipStream_t st);
template <class T> void launch_hydrostatic(const Ctx<T> &C, hipStream_t st);
template <class T> void launch_rk3(const Ctx<T> &C, T dt, T gamma, T zeta, int first, int nfields, hipStream_t st);
// U −= dtp·∇p (phi = 0: p is the pressure) or U −= ∇ϕ (phi = 1: ϕ = p·Δt, as the reference stores it)
// wrap = 1: periodic x and y without an immersed bottom; p's halos are not read
template <class T> void launch_correct(const Ctx<T> &C, T dtp, int phi, hipStream_t st, int wrap = 0);
// mask_immersed_field! of a field at location loc (0 ccc, 1 fcc, 2 cfc, 3 ccf) on an immersed grid
template <class T> void launch_mask(const DevGrid<T> &g, T *f, int loc, hipStream_t st);
// conjugate-gradient Poisson solver kernels (om_cg.hip)
// Grid-wide reductions of the elementwise kernels use kCgRedBlocks blocks (and partial sums per
// value).  The 7-point operator sweeps take a 64 × 4 tile of columns and a chunk of kchunk levels
// per block, with the chunk length chosen so that the grid holds at most kCgStencilPartials blocks
// where the tiles allow it; cg_partials_needed is the partial-sum buffer a solver allocates.
constexpr int kCgRedBlocks = 2048;
constexpr int kCgStTx = 64, kCgStTy = 4;
constexpr int kCgStencilPartials = 4096;
constexpr int kCgStBatch = 4;      // levels per batch of loads in a sweep
constexpr int kCgStMinChunk = 32;  // levels per thread where the grid allows it
struct CgStencilShape {
    int ntx, nty, kchunk, nblocks;
};
inline CgStencilShape cg_stencil_shape(int Nx, int Ny, int Nz) {
    CgStencilShape s;
    s.ntx = (Nx + kCgStTx - 1) / kCgStTx;
    s.nty = (Ny + kCgStTy - 1) / kCgStTy;
    const int tiles = s.ntx * s.nty;
    int nch = std::max(1, std::min((Nz + kCgStMinChunk - 1) / kCgStMinChunk, kCgStencilPartials / tiles));
    s.kchunk = (Nz + nch - 1) / nch;
    nch = (Nz + s.kchunk - 1) / s.kchunk;
    s.nblocks = tiles * nch;
    return s;
}
inline int cg_partials_needed(int Nx, int Ny, int Nz) {
    return std::max(4 * kCgRedBlocks, 3 * cg_stencil_shape(Nx, Ny, Nz).nblocks);
}
// The iteration's scalars stay on the device, in the solver's reduction-output array: slots 0-3
// hold results the host reads; ρ of iteration k (and, for flexible CG, z·r_prev after it) sit at
// cg_slot_rho(k) by the parity of k, so that ρ_prev is still there; then p·Ap with the sums of p
// and A·p, and the gauge sweep's ‖V⁻¹r‖² with the sums of x and r.  The update kernels form α, β and
// the gauge means from them.
constexpr int kCgSlotRho = 8, kCgSlotPQ = 12, kCgSlotGauge = 16, kCgSlots = 24;   // PQ: p·Ap, Σp, Σq; Gauge: ‖V⁻¹r‖², Σx, Σr
__host__ __device__ inline int cg_slot_rho(int it) { return kCgSlotRho + 2 * (it & 1); }
template <class T> void cg_scale(const DevGrid<T> &g, T *f, T s, int mode, hipStream_t st);
template <class T> void cg_gather(const DevGrid<T> &g, const double *x, T *dense, hipStream_t st);
template <class T> void cg_xio(const DevGrid<T> &g, T *f, double *x, T s, int store, hipStream_t st);
template <class T> void cg_source(const Ctx<T> &C, const T *dense_rhs, double *b, hipStream_t st);
template <class T> void cg_residual(const DevGrid<T> &g, const double *x, const double *b, double *r, double *partial,
                                    double *out, hipStream_t st);
template <class T> void cg_halo_r(const DevGrid<T> &g, double *r, hipStream_t st);
template <class T> void cg_dd(const DevGrid<T> &g, const double *r, T *z, hipStream_t st);
template <class T> void cg_copy_r(const DevGrid<T> &g, const double *r, T *z, hipStream_t st);
// learned correction to the FFT preconditioner and flexible CG (om_cg.hip)
template <class T> void ln_scales(const DevGrid<T> &g, const double *r, const T *z, double *partial, double *out, hipStream_t st);
template <class T> void ln_input(const DevGrid<T> &g, const double *r, const T *z, double isr, double isz, float *in3, hipStream_t st);
template <class T> void ln_x0(const DevGrid<T> &g, const double *r, const T *z, double isr, double isz, void *x0, hipStream_t st);
template <class T> void ln_correct(const DevGrid<T> &g, T *z, const float *out, double sz, hipStream_t st);
template <class T> void cg_copy_d(const DevGrid<T> &g, const double *src, double *dst, hipStream_t st);
template <class T> void cg_dot2(const DevGrid<T> &g, const T *z, const double *r, const double *rp, double *partial, double *out,
                                hipStream_t st);
template <class T> void ln_target(const DevGrid<T> &g, const double *xf, const double *xk, const float *in3k, double isz, float *tgt,
                                  double *partial, double *out, hipStream_t st);
template <class T> void ln_center(const DevGrid<T> &g, float *tgt, float mean, hipStream_t st);
template <class T> void ln_mask(const DevGrid<T> &g, float *mask, hipStream_t st);
// A-energy loss dᵀ(−V∇²)d of d = act·(y − t) (y may be null: d = −t) and, when gout is given,
// gout = −2·gscale·act·V∇²d; dscratch is a padded FP64 field
template <class T> void ln_energy(const DevGrid<T> &g, const float *y, const float *tgt, double *dscratch, double gscale, float *gout,
                                  double *partial, double *out, hipStream_t st);
// ColumnwiseTridiagonalPreconditioner: z = Lz⁻¹ r per column; t is an Nx·Ny·Nz scratch array
template <class T> void cg_ct(const DevGrid<T> &g, const double *r, T *z, T *t, hipStream_t st);
template <class T> void cg_dot_zr(const DevGrid<T> &g, const T *z, const double *r, double *partial, double *out, hipStream_t st);
template <class T> void cg_update_p(const DevGrid<T> &g, double *p, const T *z, const double *sc, int it, int flex, hipStream_t st);
template <class T> void cg_apply(const DevGrid<T> &g, const double *p, double *q, double *partial, double *out, hipStream_t st);
// The gauge and norm sweep (step = 0: the solve's first; step = 1: with the iteration's update of
// x and r folded in); writes ‖V⁻¹r‖², Σx, Σr to out.  dense: also write r V⁻¹ in FP32 cell order,
// the FFT preconditioner's source (may be null).
template <class T> void cg_gauge_norm(const DevGrid<T> &g, double *x, const double *p, double *r, const double *q,
                                      const double *sc, int it, int step, double ncell, T *dense, double *partial, double *out,
                                      hipStream_t st);
// (the p update also fills p's one-layer halos)
template <class T> void cg_true_residual(const DevGrid<T> &g, const double *x, const double *b, double *e, double *partial,
                                         double *out, hipStream_t st);
template <class T> void cg_gauged_norm(const DevGrid<T> &g, const double *e, double mean, double *partial, double *out,
                                       hipStream_t st);
template <class T> void launch_divergence(const Ctx<T> &C, T *out, hipStream_t st);
template <class T> void launch_halos(const DevGrid<T> &g, const HaloJob<T> &J, hipStream_t st);
// Solve ∇²ϕ = source; writes ϕ / divisor into `out` (the nonhydrostatic pressure p/Δt when
// called from the pressure correction, divisor = 1 otherwise).
// elements per thread of the X-forward pass's halo-free source path
constexpr int kXFwdFastElems = 8;
// whether the X-forward pass forms ∇·U of the state without reading a halo (its fast path)
template <class T> bool poisson_source_reads_no_halo(const PoissonPlan<T> &P, const Ctx<T> &C);
template <class T> void launch_poisson(const PoissonPlan<T> &P, const Ctx<T> &C, const T *src, T *out, int dense_out,
                                       T divisor, hipStream_t st);

// FP64 iterative refinement of a Poisson solve.  Residual r = b − ∇²p in FP64 arithmetic, stored
// dense in T: ∇² is the operator the FFT / Fourier–tridiagonal solver inverts (periodic x, y;
// Neumann walls on a Bounded z; ihx2 etc. the FP64 inverse squared spacings of its eigenvalues),
// evaluated without halos; p has element strides psy, psz.  b: b64 (dense FP64) if given, else the
// FP64 divergence of the velocities of C divided by `divisor`.  Then p += d (dense correction).
struct RefineOp {
    double ihx2, ihy2, ihz2;   // ihz2 < 0: stretched z, coefficients from idzf, idzc
    const double *idzf;        // FP64 1/Δzᶠ at faces k = 0 .. Nz and 1/Δzᶜ of cells k = 0 .. Nz − 1,
    const double *idzc;        // from the configuration's FP64 face heights (stretched z)
};
template <class T> void launch_poisson_residual(const Ctx<T> &C, const RefineOp &L, const T *p, int64_t psy, int64_t psz,
                                                const double *b64, double divisor, T *r, hipStream_t st);
template <class T> void launch_add_correction(const DevGrid<T> &g, T *p, int64_t psy, int64_t psz, const T *d, hipStream_t st);

// diagnostics reductions (om_diagnostics.hip); partial holds 3·1024 doubles, out 3 doubles
// (max, min, NaN count); the advection timescale reduction leaves max = max inverse timescale
template <class T> void launch_adv_timescale(const Ctx<T> &C, double *partial, double *out, hipStream_t st);
// allocation-free reductions: the device address of the library's static scratch array on `device`
// (taken once per device; nullptr with *err set on failure), its size and the part the partials take
double *diag_scratch(int device, int *err);
int diag_scratch_doubles();
int diag_partial_doubles();
constexpr int64_t kNoallocMaxElements = int64_t(1) << 40;
template <class T> void launch_field_stats(const DevGrid<T> &g, const T *origin, int with_halos, int face_z, double *partial,
                                           double *out, hipStream_t st);

}  // namespace om
