// om_oracle.cpp — CPU restatement of the Oceananigans NonhydrostaticModel RK3 time step.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP library is checked against; the
// product never links, loads or calls it.  It follows the reference's own structure
// (operator by operator, field by field, one kernel per field per phase) rather than the
// fused MI355X design, so that a mistake in one is not silently shared by the other.
// Paths are relative to the Oceananigans.jl source tree (version 0.110.19).
//
//   grid metrics           src/Grids/grid_generation.jl:34-156, src/Operators/spacings_and_areas_and_volumes.jl
//   halo fills             src/BoundaryConditions/fill_halo_regions*.jl, boundary_condition_ordering.jl
//   WENO / centered        src/Advection/weno_interpolants.jl, reconstruction_coefficients.jl,
//                          topologically_conditional_interpolation.jl, upwind_biased_advective_fluxes.jl,
//                          centered_advective_fluxes.jl, momentum/tracer_advection_operators.jl
//   closures               src/TurbulenceClosures/turbulence_closure_implementations/anisotropic_minimum_dissipation.jl,
//                          .../Smagorinskys/*.jl, abstract_scalar_diffusivity_closure.jl,
//                          closure_kernel_operators.jl, velocity_tracer_gradients.jl
//   buoyancy / coriolis    src/BuoyancyFormulations/*.jl, src/Coriolis/f_plane.jl, coriolis_schemes.jl
//   tendencies             src/Models/NonhydrostaticModels/nonhydrostatic_tendency_kernel_functions.jl
//   pressure               src/Solvers/fft_based_poisson_solver.jl, poisson_eigenvalues.jl,
//                          fourier_tridiagonal_poisson_solver.jl, batched_tridiagonal_solver.jl,
//                          src/Models/NonhydrostaticModels/pressure_correction.jl, solve_for_pressure.jl
//   conjugate gradient     src/Solvers/conjugate_gradient_poisson_solver.jl, conjugate_gradient_solver.jl
//   immersed bottom        src/ImmersedBoundaries/grid_fitted_bottom.jl, immersed_boundary_interface.jl,
//                          conditional_differences.jl, mask_immersed_field.jl, src/Grids/inactive_node.jl,
//                          src/Advection/immersed_advective_fluxes.jl, src/TurbulenceClosures/immersed_diffusive_fluxes.jl,
//                          src/Coriolis/coriolis_schemes.jl
//   time stepping          src/TimeSteppers/runge_kutta_3.jl, nonhydrostatic_rk3_substep.jl
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/oceananigans_mi355x.h"

// Planted mistakes, for tests that show a pin fails on a plausible error (built only as the
// separate oracle/liboracle_plant<N>.so libraries of those tests; 0 in every other build):
//   1  one AMD interpolation misplaced: the ν·Σ₂₂ term's ffc quantity averaged with ℑxz instead of ℑxy
//   2  pHY′ started at the top with Δzᶠ(Nz) instead of Δzᶠ(Nz+1)
//   3  the z buffer band of WENO5 one face too wide at the bottom (WENO5 at face 3)
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

namespace orc {

enum { P = OM_PERIODIC, B = OM_BOUNDED, F = OM_FLAT };

// ----------------------------------------------------------------------------------------
// Fields: every field is stored with the same padded parent shape (N + 2H + 1 per axis) so
// Center and Face fields share indexing; reference indices i ∈ 1-H : N+H(+1).
// ----------------------------------------------------------------------------------------
template <class T> struct Fld {
    int nx = 0, ny = 0, nz = 0, ox = 0, oy = 0, oz = 0;
    std::vector<T> d;
    void init(const int N[3], const int H[3]) {
        nx = N[0] + 2 * H[0] + 1; ny = N[1] + 2 * H[1] + 1; nz = N[2] + 2 * H[2] + 1;
        ox = H[0] - 1; oy = H[1] - 1; oz = H[2] - 1;
        d.assign((size_t)nx * ny * nz, T(0));
    }
    inline T &operator()(int i, int j, int k) {
        return d[(size_t)(i + ox) + (size_t)nx * ((size_t)(j + oy) + (size_t)ny * (size_t)(k + oz))];
    }
    inline T operator()(int i, int j, int k) const {
        return d[(size_t)(i + ox) + (size_t)nx * ((size_t)(j + oy) + (size_t)ny * (size_t)(k + oz))];
    }
};

// ----------------------------------------------------------------------------------------
// Grid (RectilinearGrid with uniform x, y and uniform-or-stretched z)
// src/Grids/grid_generation.jl:34-95 (stretched), :117-156 (uniform), :159-176 (Flat)
// ----------------------------------------------------------------------------------------
template <class T> struct Grid {
    int N[3], H[3], topo[3];
    T dx, dy;                       // Δxᶜ = Δxᶠ, Δyᶜ = Δyᶠ
    std::vector<T> dzc_, dzf_;      // indexed by k + Hz + 1 (k from -Hz .. Nz+Hz+1)
    std::vector<double> zc_, zf_;   // node positions (for initial conditions), same indexing
    double origin[3], L[3];
    bool stretched = false;
    T Lx, Ly, Lz;
    inline T dzC(int k) const { return dzc_[k + H[2] + 1]; }
    inline T dzF(int k) const { return dzf_[k + H[2] + 1]; }
    double xC(int i) const { return origin[0] + (i - 0.5) * (L[0] / N[0]); }
    double yC(int j) const { return origin[1] + (j - 0.5) * (L[1] / N[1]); }
    double xF(int i) const { return origin[0] + (i - 1) * (L[0] / N[0]); }
    double yF(int j) const { return origin[1] + (j - 1) * (L[1] / N[1]); }
    double zC(int k) const { return zc_[k + H[2] + 1]; }
    double zF(int k) const { return zf_[k + H[2] + 1]; }
};

template <class T> static T uniform_spacing(double L, int N) {
    // Δ = L / N evaluated in extended precision, then rounded once to FT (grid_generation.jl:126-131)
    long double d = (long double)L / (long double)N;
    return (T)d;
}

template <class T> static void build_grid(Grid<T> &g, const om_config_t *c) {
    for (int d = 0; d < 3; ++d) {
        g.N[d] = c->N[d]; g.H[d] = c->topology[d] == F ? 0 : c->H[d]; g.topo[d] = c->topology[d];
        g.origin[d] = c->origin[d]; g.L[d] = c->extent[d];
    }
    g.dx = g.topo[0] == F ? T(1) : uniform_spacing<T>(c->extent[0], c->N[0]);
    g.dy = g.topo[1] == F ? T(1) : uniform_spacing<T>(c->extent[1], c->N[1]);
    g.Lx = g.topo[0] == F ? T(1) : (T)c->extent[0];
    g.Ly = g.topo[1] == F ? T(1) : (T)c->extent[1];
    const int Nz = g.N[2], Hz = g.H[2];
    const int n = Nz + 2 * Hz + 3;
    g.dzc_.assign(n, T(1)); g.dzf_.assign(n, T(1)); g.zc_.assign(n, c->origin[2]); g.zf_.assign(n, c->origin[2]);
    if (g.topo[2] == F) { g.Lz = T(1); return; }
    if (!c->z_faces) {
        T dz = uniform_spacing<T>(c->extent[2], Nz);
        for (int idx = 0; idx < n; ++idx) { g.dzc_[idx] = dz; g.dzf_[idx] = dz; }
        long double d = (long double)c->extent[2] / Nz;
        for (int k = -Hz; k <= Nz + Hz + 1; ++k) {
            g.zf_[k + Hz + 1] = (double)((long double)c->origin[2] + (k - 1) * d);
            g.zc_[k + Hz + 1] = (double)((long double)c->origin[2] + (k - 0.5L) * d);
        }
        g.Lz = (T)c->extent[2];
        return;
    }
    // Stretched (FaceExtendedTopology for Bounded; Periodic wraps spacings)
    g.stretched = true;
    std::vector<T> Fi(Nz + 1);
    for (int i = 0; i <= Nz; ++i) Fi[i] = (T)c->z_faces[i];
    std::vector<T> dlo(Hz), dhi(Hz);
    for (int i = 0; i < Hz; ++i) {
        if (g.topo[2] == B) { dlo[i] = Fi[1] - Fi[0]; dhi[i] = Fi[Nz] - Fi[Nz - 1]; }
        else { dlo[i] = Fi[Nz - Hz + i + 1] - Fi[Nz - Hz + i]; dhi[i] = Fi[i + 1] - Fi[i]; }
    }
    std::reverse(dhi.begin(), dhi.end());
    std::vector<T> Fall;  // faces 1-Hz .. Nz+1+Hz
    for (int i = 0; i < Hz; ++i) { T s = 0; for (int q = i; q < Hz; ++q) s += dlo[q]; Fall.push_back(Fi[0] - s); }
    for (int i = 0; i <= Nz; ++i) Fall.push_back(Fi[i]);
    std::vector<T> Fp;
    for (int i = 0; i < Hz; ++i) { T s = 0; for (int q = i; q < Hz; ++q) s += dhi[q]; Fp.push_back(Fi[Nz] + s); }
    std::reverse(Fp.begin(), Fp.end());
    for (auto v : Fp) Fall.push_back(v);
    // Fall[m] is face index m + 1 - Hz
    const int TC = Nz + 2 * Hz;
    std::vector<T> C(TC);
    for (int i = 0; i < TC; ++i) C[i] = (Fall[i + 1] + Fall[i]) / T(2);
    for (int k = 1 - Hz; k <= Nz + Hz; ++k) g.dzc_[k + Hz + 1] = Fall[k + Hz] - Fall[k + Hz - 1];
    // Δᶠ[k] = C[k] - C[k-1]; Δᶠ[1-H] = Δᶠ[-H] = Δᶠ[2-H]
    for (int k = 2 - Hz; k <= Nz + Hz; ++k) g.dzf_[k + Hz + 1] = C[k + Hz - 1] - C[k + Hz - 2];
    g.dzf_[1 - Hz + Hz + 1] = g.dzf_[2 - Hz + Hz + 1];
    g.dzf_[-Hz + Hz + 1] = g.dzf_[2 - Hz + Hz + 1];
    // Bounded face fields have a face at Nz+1+Hz: Δᶠ beyond the last center is never used
    g.dzf_[Nz + Hz + 1 + Hz + 1] = g.dzf_[Nz + Hz + Hz + 1];
    g.dzc_[-Hz + Hz + 1] = g.dzc_[1 - Hz + Hz + 1];
    g.dzc_[Nz + Hz + 1 + Hz + 1] = g.dzc_[Nz + Hz + Hz + 1];
    for (int k = 1 - Hz; k <= Nz + 1 + Hz; ++k) g.zf_[k + Hz + 1] = (double)Fall[k + Hz - 1];
    for (int k = 1 - Hz; k <= Nz + Hz; ++k) g.zc_[k + Hz + 1] = (double)C[k + Hz - 1];
    g.Lz = Fi[Nz] - Fi[0];
}

// ----------------------------------------------------------------------------------------
// Reconstruction coefficients (reconstruction_coefficients.jl:62-77, weno_interpolants.jl:76-140)
// coefficient tuples are FT-rounded with the last entry = 1 - sum(others) in FT.
// ----------------------------------------------------------------------------------------
template <class T> struct Coeffs {
    T w5[3][3];   // WENO{3} stencil r: coefficients on (ψᵣ[1], ψᵣ[2], ψᵣ[3])
    T w3[2][2];   // WENO{2}
    T c4[4];      // Centered{2}, on offsets (-2, -1, 0, +1) from the face
    T c2[2];      // Centered{1}, on offsets (-1, 0)
    T Cs5[3], Cs3[2];
    T eps;
    Coeffs() {
        const double r5[3][2] = {{1.0 / 3, 5.0 / 6}, {-1.0 / 6, 5.0 / 6}, {1.0 / 3, -7.0 / 6}};
        for (int r = 0; r < 3; ++r) {
            T a = (T)r5[r][0], b = (T)r5[r][1];
            w5[r][0] = a; w5[r][1] = b; w5[r][2] = T(1) - (a + b);
        }
        const double r3[2] = {0.5, -0.5};
        for (int r = 0; r < 2; ++r) { T a = (T)r3[r]; w3[r][0] = a; w3[r][1] = T(1) - a; }
        T m = (T)(-1.0 / 12), s = (T)(7.0 / 12);
        T c1 = m, c2_ = s, c3 = s, c4_ = T(1) - ((c1 + c2_) + c3);
        c4[0] = c4_; c4[1] = c3; c4[2] = c2_; c4[3] = c1;
        c2[0] = T(1) - T(0.5); c2[1] = T(0.5);
        Cs5[0] = (T)(3.0 / 10); Cs5[1] = (T)(3.0 / 5); Cs5[2] = (T)(1.0 / 10);
        Cs3[0] = (T)(2.0 / 3); Cs3[1] = (T)(1.0 / 3);
        eps = (T)1e-8f;   // ϵ = 1f-8 (a Float32 literal) promoted to FT
    }
};

template <class T> static inline bool is_fp32() { return sizeof(T) == 4; }

// The reference wraps the smoothness indicators, the WENO weighted sum and the Centered
// interpolants in @muladd (weno_interpolants.jl:265,279,516; centered_reconstruction.jl:56-62),
// which Julia lowers to fused multiply-adds on FMA hardware.  MuladdMacro 0.2 keeps Julia's
// left-to-right evaluation of a sum of products: p₁ + p₂ + … + pₙ becomes
// muladd(pₙ, muladd(…, muladd(p₂, p₁))) — the first product is rounded on its own and every
// later one is fused into the running sum — and a product of three factors a·b·c is split as
// (a·b, c).  Everything outside these sites is evaluated without contraction.

// β for WENO{3} (weno_interpolants.jl:172-174, 228-283): with C = (C₁..C₆) the reference's
// unrolled form ψ₁(C₁ψ₁ + C₂ψ₂ + C₃ψ₃) + ψ₂(C₄ψ₂ + C₅ψ₃) + ψ₃ψ₃C₆ under @muladd; FP32 first
// subtracts the stencil's middle value ψ̂ = ψ₂ from every entry.
template <class T> static inline T beta5(const T p[3], int r) {
    static const double C[3][6] = {{10, -31, 11, 25, -19, 4}, {4, -13, 5, 13, -13, 4}, {4, -19, 11, 25, -31, 10}};
    T a = p[0], b = p[1], c = p[2];
    if (is_fp32<T>()) { T h = p[1]; a -= h; b -= h; c -= h; }
    const double *K = C[r];
    const T s1 = std::fma((T)K[2], c, std::fma((T)K[1], b, (T)K[0] * a));
    const T s2 = std::fma((T)K[4], c, (T)K[3] * b);
    return std::fma(c * c, (T)K[5], std::fma(b, s2, a * s1));
}
// β for WENO{2}: ψ₁(C₁ψ₁ + C₂ψ₂) + ψ₂ψ₂C₃ with C = (1, −2, 1)
template <class T> static inline T beta3(const T p[2]) {
    T a = p[0], b = p[1];
    if (is_fp32<T>()) { T h = p[1]; a -= h; b -= h; }
    return std::fma(b * b, T(1), a * std::fma(T(-2), b, T(1) * a));
}

// WENO(order=5) biased face value from S = (ψ[i-3], ..., ψ[i+2]); left = (ũ > 0)
template <class T> static T weno5(const Coeffs<T> &K, const T S[6], bool left) {
    T ps[3][3];
    if (left) {
        ps[0][0] = S[2]; ps[0][1] = S[3]; ps[0][2] = S[4];
        ps[1][0] = S[1]; ps[1][1] = S[2]; ps[1][2] = S[3];
        ps[2][0] = S[0]; ps[2][1] = S[1]; ps[2][2] = S[2];
    } else {
        ps[0][0] = S[3]; ps[0][1] = S[2]; ps[0][2] = S[1];
        ps[1][0] = S[4]; ps[1][1] = S[3]; ps[1][2] = S[2];
        ps[2][0] = S[5]; ps[2][1] = S[4]; ps[2][2] = S[3];
    }
    T beta[3], alpha[3];
    for (int r = 0; r < 3; ++r) beta[r] = beta5<T>(ps[r], r);
    T tau = std::abs(beta[0] - beta[2]);
    for (int r = 0; r < 3; ++r) { T q = tau / (beta[r] + K.eps); alpha[r] = K.Cs5[r] * (T(1) + q * q); }
    T sinv = T(1) / ((alpha[0] + alpha[1]) + alpha[2]);
    // ω = α .* Σα⁻¹; candidate values sum(coeff .* ψ) (no @muladd); the weighted sum under @muladd
    T w[3], p[3];
    for (int r = 0; r < 3; ++r) {
        w[r] = alpha[r] * sinv;
        p[r] = K.w5[r][0] * ps[r][0] + K.w5[r][1] * ps[r][1] + K.w5[r][2] * ps[r][2];
    }
    return std::fma(w[2], p[2], std::fma(w[1], p[1], w[0] * p[0]));
}

// WENO(order=3) biased face value from S = (ψ[i-2], ψ[i-1], ψ[i], ψ[i+1])
template <class T> static T weno3(const Coeffs<T> &K, const T S[4], bool left) {
    T ps[2][2];
    if (left) { ps[0][0] = S[1]; ps[0][1] = S[2]; ps[1][0] = S[0]; ps[1][1] = S[1]; }
    else      { ps[0][0] = S[2]; ps[0][1] = S[1]; ps[1][0] = S[3]; ps[1][1] = S[2]; }
    T beta[2] = {beta3<T>(ps[0]), beta3<T>(ps[1])};
    T tau = std::abs(beta[0] - beta[1]);
    T alpha[2];
    for (int r = 0; r < 2; ++r) { T q = tau / (beta[r] + K.eps); alpha[r] = K.Cs3[r] * (T(1) + q * q); }
    T sinv = T(1) / (alpha[0] + alpha[1]);
    const T p0 = K.w3[0][0] * ps[0][0] + K.w3[0][1] * ps[0][1], p1 = K.w3[1][0] * ps[1][0] + K.w3[1][1] * ps[1][1];
    return std::fma(alpha[1] * sinv, p1, (alpha[0] * sinv) * p0);
}

// ----------------------------------------------------------------------------------------
// The model state
// ----------------------------------------------------------------------------------------
template <class T> struct Model {
    om_config_t cfg;
    std::vector<double> zfaces_copy;
    Grid<T> g;
    Coeffs<T> K;
    int nt = 0;
    Fld<T> U[3 + OM_MAX_TRACERS];      // u, v, w, tracers
    Fld<T> Gn[3 + OM_MAX_TRACERS], Gm[3 + OM_MAX_TRACERS];
    Fld<T> pNHS, pHY, nu, kap[OM_MAX_TRACERS], rhs;
    Fld<T> cg_r, cg_p, cg_q, cg_z, cg_b;    // conjugate-gradient work fields
    int cg_iterations = 0;
    double cg_residual = 0, cg_true_residual = 0;
    bool has_pHY = false;
    // GridFittedBottom: kbot_[(i-1) + Nx (j-1)] = number of immersed cells at the bottom of column (i, j)
    bool ib = false;
    std::vector<int> kbot_;
    double time = 0; int64_t iteration = 0; int stage = 1; double last_dt = INFINITY, last_stage_dt = INFINITY;
    std::string err;
    int nthreads = 1;

    int Nx() const { return g.N[0]; }
    int Ny() const { return g.N[1]; }
    int Nz() const { return g.N[2]; }
    bool flatz() const { return g.topo[2] == F; }
    bool bz() const { return g.topo[2] == B; }

    // ---------------- metric operators (spacings_and_areas_and_volumes.jl:122-129, 418-490)
    T Ax_c(int k) const { return g.dy * g.dzC(k); }   // Ax at z-Center (fcc, ccc, ffc)
    T Ax_f(int k) const { return g.dy * g.dzF(k); }   // Ax at z-Face   (fcf, ccf)
    T Ay_c(int k) const { return g.dx * g.dzC(k); }
    T Ay_f(int k) const { return g.dx * g.dzF(k); }
    T Az() const { return g.dx * g.dy; }
    T V_c(int k) const { return (g.dx * g.dy) * g.dzC(k); }
    T V_f(int k) const { return (g.dx * g.dy) * g.dzF(k); }

    int kbot(int i, int j) const {   // halo columns take the periodic images (fill_halo_regions!(bottom_field))
        const int Nx = g.N[0], Ny = g.N[1];
        const int ii = ((i - 1) % Nx + Nx) % Nx, jj = ((j - 1) % Ny + Ny) % Ny;
        return kbot_[(size_t)ii + (size_t)Nx * jj];
    }
};

// ----------------------------------------------------------------------------------------
// Immersed-boundary predicates (src/Grids/inactive_node.jl:36-137, immersed_boundary_interface.jl)
// Locations: lx, ly, lz = 0 (Center) or 1 (Face).  `under` evaluates the predicate on the
// underlying grid (only cells outside a Bounded z are inactive there).
// ----------------------------------------------------------------------------------------
template <class T> struct IBP {
    const Model<T> &m;
    explicit IBP(const Model<T> &mm) : m(mm) {}
    bool outside(int k) const { return m.bz() && (k < 1 || k > m.Nz()); }
    // _immersed_cell: zᶜ(k) ≤ numerical bottom ⟺ k ≤ number of immersed cells (grid_fitted_bottom.jl:141-147)
    bool immersed(int i, int j, int k) const { return m.ib && k <= m.kbot(i, j); }
    bool inactive_cell(int i, int j, int k, bool under) const { return (!under && immersed(i, j, k)) || outside(k); }
    bool inactive_node(int i, int j, int k, int lx, int ly, int lz, bool under) const {
        if (lx && ly && lz) return inactive_node(i, j, k, 0, 1, 1, under) && inactive_node(i - 1, j, k, 0, 1, 1, under);
        if (lx && ly) return inactive_node(i, j, k, 0, 1, 0, under) && inactive_node(i - 1, j, k, 0, 1, 0, under);
        if (lx && lz) return inactive_node(i, j, k, 0, 0, 1, under) && inactive_node(i - 1, j, k, 0, 0, 1, under);
        if (ly && lz) return inactive_node(i, j, k, 0, 1, 0, under) && inactive_node(i, j, k - 1, 0, 1, 0, under);
        if (lx) return inactive_cell(i, j, k, under) && inactive_cell(i - 1, j, k, under);
        if (ly) return inactive_cell(i, j, k, under) && inactive_cell(i, j - 1, k, under);
        if (lz) return inactive_cell(i, j, k, under) && inactive_cell(i, j, k - 1, under);
        return inactive_cell(i, j, k, under);
    }
    bool peripheral_node(int i, int j, int k, int lx, int ly, int lz, bool under) const {
        if (lx && ly && lz) return peripheral_node(i, j, k, 0, 1, 1, under) || peripheral_node(i - 1, j, k, 0, 1, 1, under);
        if (lx && ly) return peripheral_node(i, j, k, 0, 1, 0, under) || peripheral_node(i - 1, j, k, 0, 1, 0, under);
        if (lx && lz) return peripheral_node(i, j, k, 0, 0, 1, under) || peripheral_node(i - 1, j, k, 0, 0, 1, under);
        if (ly && lz) return peripheral_node(i, j, k, 0, 1, 0, under) || peripheral_node(i, j, k - 1, 0, 1, 0, under);
        if (lx) return inactive_cell(i, j, k, under) || inactive_cell(i - 1, j, k, under);
        if (ly) return inactive_cell(i, j, k, under) || inactive_cell(i, j - 1, k, under);
        if (lz) return inactive_cell(i, j, k, under) || inactive_cell(i, j, k - 1, under);
        return inactive_cell(i, j, k, under);
    }
    bool imm_inactive(int i, int j, int k, int lx, int ly, int lz) const {
        return m.ib && inactive_node(i, j, k, lx, ly, lz, false) && !inactive_node(i, j, k, lx, ly, lz, true);
    }
    bool imm_peripheral(int i, int j, int k, int lx, int ly, int lz) const {
        return m.ib && peripheral_node(i, j, k, lx, ly, lz, false) && !peripheral_node(i, j, k, lx, ly, lz, true);
    }
    // near_{x,y,z}_immersed_boundary_{symmetric,biased}{ᶜ,ᶠ} for a scheme with `buffer`
    // (immersed_advective_fluxes.jl:127-215): any inactive node among the 2·buffer stencil points;
    // the node is a cell centre when the interpolation point is a face and a face when it is a centre.
    bool near(int d, bool center_side, int buffer, int i, int j, int k) const {
        for (int n = 1; n <= 2 * buffer; ++n) {
            const int c = center_side ? n - buffer : n - buffer - 1;
            const int l = center_side ? 1 : 0;
            bool in;
            if (d == 0) in = inactive_node(i + c, j, k, l, 0, 0, false);
            else if (d == 1) in = inactive_node(i, j + c, k, 0, l, 0, false);
            else in = inactive_node(i, j, k + c, 0, 0, l, false);
            if (in) return true;
        }
        return false;
    }
};

// ----------------------------------------------------------------------------------------
// Halo fills (fill_halo_regions.jl, fill_halo_regions_{flux,periodic,value_gradient,normal_flow}.jl)
// Order: Flux/Value/Gradient/NormalFlow (z) first, then Periodic x, then Periodic y (and z).
// ----------------------------------------------------------------------------------------
template <class T> static void fill_periodic(Model<T> &m, Fld<T> &c, int dim) {
    const int *N = m.g.N, *H = m.g.H;
    if (m.g.topo[dim] != P || H[dim] == 0) return;
    // parent ranges (whole padded extent of the other dims)
    int ilo = 1 - H[0], ihi = N[0] + H[0], jlo = 1 - H[1], jhi = N[1] + H[1], klo = 1 - H[2], khi = N[2] + H[2] + 1;
    if (m.g.topo[0] == F) { ilo = ihi = 1; }
    if (m.g.topo[1] == F) { jlo = jhi = 1; }
    if (m.g.topo[2] == F) { klo = khi = 1; }
    if (dim == 0) {
        for (int k = klo; k <= khi; ++k) for (int j = jlo; j <= jhi; ++j)
            for (int h = 1; h <= H[0]; ++h) { c(1 - h, j, k) = c(N[0] + 1 - h, j, k); c(N[0] + h, j, k) = c(h, j, k); }
    } else if (dim == 1) {
        for (int k = klo; k <= khi; ++k) for (int i = ilo; i <= ihi; ++i)
            for (int h = 1; h <= H[1]; ++h) { c(i, 1 - h, k) = c(i, N[1] + 1 - h, k); c(i, N[1] + h, k) = c(i, h, k); }
    } else {
        for (int j = jlo; j <= jhi; ++j) for (int i = ilo; i <= ihi; ++i)
            for (int h = 1; h <= H[2]; ++h) { c(i, j, 1 - h) = c(i, j, N[2] + 1 - h); c(i, j, N[2] + h) = c(i, j, h); }
    }
}

// z-boundary fill for a Center field (Flux: one-layer copy; Value/Gradient: linear extrapolation)
template <class T> static void fill_bounded_z_center(Model<T> &m, Fld<T> &c, int bkind, double bval, int tkind, double tval) {
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
    for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
        // bottom
        if (bkind == OM_BC_VALUE || bkind == OM_BC_GRADIENT) {
            T D = m.g.dzF(1); T c1 = c(i, j, 1);
            T grad = bkind == OM_BC_GRADIENT ? (T)bval : (c1 - (T)bval) / (D / T(2));
            c(i, j, 0) = c1 + grad * (-D);
        } else c(i, j, 0) = c(i, j, 1);
        // top
        if (tkind == OM_BC_VALUE || tkind == OM_BC_GRADIENT) {
            T D = m.g.dzF(Nz + 1); T cN = c(i, j, Nz);
            T grad = tkind == OM_BC_GRADIENT ? (T)tval : ((T)tval - cN) / (D / T(2));
            c(i, j, Nz + 1) = cN + grad * D;
        } else c(i, j, Nz + 1) = c(i, j, Nz);
    }
}

// Fill halos of prognostic field f (0=u, 1=v, 2=w, 3+ tracers).
template <class T> static void fill_prognostic_halos(Model<T> &m, int f, bool fill_normal) {
    Fld<T> &c = m.U[f];
    if (m.bz()) {
        if (f == 2) {
            if (fill_normal) {
                for (int j = 1; j <= m.Ny(); ++j) for (int i = 1; i <= m.Nx(); ++i) { c(i, j, 1) = 0; c(i, j, m.Nz() + 1) = 0; }
            }
        } else {
            fill_bounded_z_center(m, c, m.cfg.bc_bottom[f], m.cfg.bc_bottom_value[f], m.cfg.bc_top[f], m.cfg.bc_top_value[f]);
        }
    }
    fill_periodic(m, c, 2);
    fill_periodic(m, c, 0);
    fill_periodic(m, c, 1);
}

// Default (NoFlux) halos of an auxiliary Center field (closure fields, pressures).
template <class T> static void fill_aux_halos(Model<T> &m, Fld<T> &c) {
    if (m.bz()) fill_bounded_z_center(m, c, OM_BC_DEFAULT, 0, OM_BC_DEFAULT, 0);
    fill_periodic(m, c, 2);
    fill_periodic(m, c, 0);
    fill_periodic(m, c, 1);
}

// ----------------------------------------------------------------------------------------
// Topologically conditional interpolation in z (topologically_conditional_interpolation.jl:52-58,106-124)
// x and y are Periodic (no switching) or Flat.
// ----------------------------------------------------------------------------------------
enum ZScheme { Z5, Z3, Z2, Z4 };
// biased reconstruction at a z-Face k
template <class T> static inline ZScheme zb_face(const Model<T> &m, int k) {
    if (!m.bz()) return Z5;
    const int Nz = m.Nz();
    if (k >= (ORC_PLANT == 3 ? 3 : 4) && k <= Nz - 2) return Z5;
    if (k >= 3 && k <= Nz - 1) return Z3;
    return Z2;
}
// biased reconstruction at a z-Center k (evaluated as the face k+1 stencil)
template <class T> static inline ZScheme zb_center(const Model<T> &m, int k) {
    if (!m.bz()) return Z5;
    const int Nz = m.Nz();
    if (k >= 3 && k <= Nz - 2) return Z5;
    if (k >= 2 && k <= Nz - 1) return Z3;
    return Z2;
}
// symmetric (advecting velocity) interpolation: Centered(4) or Centered(2)
template <class T> static inline ZScheme zs_face(const Model<T> &m, int k) {
    if (!m.bz()) return Z4;
    return (k >= 4 && k <= m.Nz() - 2) ? Z4 : Z2;
}
template <class T> static inline ZScheme zs_center(const Model<T> &m, int k) {
    if (!m.bz()) return Z4;
    return (k >= 3 && k <= m.Nz() - 2) ? Z4 : Z2;
}

// Generic 1D reconstructions along a direction, where `q(o)` returns the value at offset o
// from the face index.  Face f sits between cells f-1 and f.
template <class T, class Q> static inline T rec_biased(const Coeffs<T> &K, ZScheme s, bool left, Q q) {
    if (s == Z5) { T S[6] = {q(-3), q(-2), q(-1), q(0), q(1), q(2)}; return weno5(K, S, left); }
    if (s == Z3) { T S[4] = {q(-2), q(-1), q(0), q(1)}; return weno3(K, S, left); }
    return std::fma(K.c2[1], q(0), K.c2[0] * q(-1));
}
// Centered(4) / Centered(2) under @muladd (centered_reconstruction.jl:56-62)
template <class T, class Q> static inline T rec_symmetric(const Coeffs<T> &K, ZScheme s, Q q) {
    if (s == Z4) return std::fma(K.c4[3], q(1), std::fma(K.c4[2], q(0), std::fma(K.c4[1], q(-1), K.c4[0] * q(-2))));
    return std::fma(K.c2[1], q(0), K.c2[0] * q(-1));
}

template <class T> static inline bool bias_left(T u) { return u > T(0); }

// ----------------------------------------------------------------------------------------
// Advective flux divergences (momentum_advection_operators.jl:47-84, tracer_advection_operators.jl:31-35)
// ----------------------------------------------------------------------------------------
template <class T> struct Adv {
    Model<T> &m;
    explicit Adv(Model<T> &mm) : m(mm) {}
    const Fld<T> &u() const { return m.U[0]; }
    const Fld<T> &v() const { return m.U[1]; }
    const Fld<T> &w() const { return m.U[2]; }
    bool weno() const { return m.cfg.advection == OM_ADV_WENO5; }
    bool fx() const { return m.g.topo[0] == F; }
    bool fy() const { return m.g.topo[1] == F; }
    bool fz() const { return m.g.topo[2] == F; }

    // Scheme of a reconstruction along d at a face (cs = false) or a centre (cs = true): on an immersed
    // grid the buffer chain WENO5 → WENO3 → Centered(2) is driven by the inactive nodes in the stencil
    // (this also reproduces the Bounded-z wall switching); otherwise `topo` (the wall switching) is used.
    ZScheme bsch(int d, bool cs, int i, int j, int k, ZScheme topo) const {
        if (!m.ib) return topo;
        IBP<T> P(m);
        if (!P.near(d, cs, 3, i, j, k)) return Z5;
        if (!P.near(d, cs, 2, i, j, k)) return Z3;
        return Z2;
    }
    ZScheme ssch(int d, bool cs, int i, int j, int k, ZScheme topo) const {
        if (!m.ib) return topo;
        return IBP<T>(m).near(d, cs, 3, i, j, k) ? Z2 : Z4;
    }
    // conditional_flux: zero on the immersed periphery of the flux location
    bool zf(int i, int j, int k, int lx, int ly, int lz) const { return m.ib && IBP<T>(m).imm_peripheral(i, j, k, lx, ly, lz); }
    // conditional δ: zero when either differenced point is an immersed inactive node
    bool zd(int i1, int j1, int k1, int i2, int j2, int k2, int lx, int ly, int lz) const {
        if (!m.ib) return false;
        IBP<T> P(m);
        return P.imm_inactive(i1, j1, k1, lx, ly, lz) || P.imm_inactive(i2, j2, k2, lx, ly, lz);
    }

    // ---- u momentum fluxes
    T F_Uu(int i, int j, int k) {  // ccc
        if (fx()) return 0;
        const Fld<T> &U = u();
        if (!weno()) {
            T Ui = T(0.5) * (U(i, j, k) + U(i + 1, j, k));
            return m.Ax_c(k) * Ui * Ui;
        }
        if (zf(i, j, k, 0, 0, 0)) return 0;
        T Axk = m.Ax_c(k);
        T ut = rec_symmetric(m.K, ssch(0, true, i, j, k, Z4), [&](int o) { return Axk * U(i + 1 + o, j, k); });
        T ur = rec_biased(m.K, bsch(0, true, i, j, k, Z5), bias_left(ut), [&](int o) { return U(i + 1 + o, j, k); });
        return ut * ur;
    }
    T F_Vu(int i, int j, int k) {  // ffc
        if (fy()) return 0;
        const Fld<T> &U = u(), &Vv = v();
        if (!weno()) {
            T vi = fx() ? Vv(i, j, k) : T(0.5) * (Vv(i - 1, j, k) + Vv(i, j, k));
            T uj = T(0.5) * (U(i, j - 1, k) + U(i, j, k));
            return m.Ay_c(k) * vi * uj;
        }
        if (zf(i, j, k, 1, 1, 0)) return 0;
        T Ayk = m.Ay_c(k);
        T vt = fx() ? Ayk * Vv(i, j, k) : rec_symmetric(m.K, ssch(0, false, i, j, k, Z4), [&](int o) { return Ayk * Vv(i + o, j, k); });
        T ur = rec_biased(m.K, bsch(1, false, i, j, k, Z5), bias_left(vt), [&](int o) { return U(i, j + o, k); });
        return vt * ur;
    }
    T F_Wu(int i, int j, int k) {  // fcf
        if (fz()) return 0;
        const Fld<T> &U = u(), &W = w();
        if (!weno()) {
            T wi = fx() ? W(i, j, k) : T(0.5) * (W(i - 1, j, k) + W(i, j, k));
            T uk = T(0.5) * (U(i, j, k - 1) + U(i, j, k));
            return m.Az() * wi * uk;
        }
        if (zf(i, j, k, 1, 0, 1)) return 0;
        T Az = m.Az();
        T wt = fx() ? Az * W(i, j, k) : rec_symmetric(m.K, ssch(0, false, i, j, k, Z4), [&](int o) { return Az * W(i + o, j, k); });
        T ur = rec_biased(m.K, bsch(2, false, i, j, k, zb_face(m, k)), bias_left(wt), [&](int o) { return U(i, j, k + o); });
        return wt * ur;
    }
    // ---- v momentum fluxes
    T F_Uv(int i, int j, int k) {  // ffc
        if (fx()) return 0;
        const Fld<T> &U = u(), &Vv = v();
        if (!weno()) {
            T uj = fy() ? U(i, j, k) : T(0.5) * (U(i, j - 1, k) + U(i, j, k));
            T vi = T(0.5) * (Vv(i - 1, j, k) + Vv(i, j, k));
            return m.Ax_c(k) * uj * vi;
        }
        if (zf(i, j, k, 1, 1, 0)) return 0;
        T Axk = m.Ax_c(k);
        T ut = fy() ? Axk * U(i, j, k) : rec_symmetric(m.K, ssch(1, false, i, j, k, Z4), [&](int o) { return Axk * U(i, j + o, k); });
        T vr = rec_biased(m.K, bsch(0, false, i, j, k, Z5), bias_left(ut), [&](int o) { return Vv(i + o, j, k); });
        return ut * vr;
    }
    T F_Vv(int i, int j, int k) {  // ccc
        if (fy()) return 0;
        const Fld<T> &Vv = v();
        if (!weno()) {
            T Vj = T(0.5) * (Vv(i, j, k) + Vv(i, j + 1, k));
            return m.Ay_c(k) * Vj * Vj;
        }
        if (zf(i, j, k, 0, 0, 0)) return 0;
        T Ayk = m.Ay_c(k);
        T vt = rec_symmetric(m.K, ssch(1, true, i, j, k, Z4), [&](int o) { return Ayk * Vv(i, j + 1 + o, k); });
        T vr = rec_biased(m.K, bsch(1, true, i, j, k, Z5), bias_left(vt), [&](int o) { return Vv(i, j + 1 + o, k); });
        return vt * vr;
    }
    T F_Wv(int i, int j, int k) {  // cff
        if (fz()) return 0;
        const Fld<T> &Vv = v(), &W = w();
        if (!weno()) {
            T wj = fy() ? W(i, j, k) : T(0.5) * (W(i, j - 1, k) + W(i, j, k));
            T vk = T(0.5) * (Vv(i, j, k - 1) + Vv(i, j, k));
            return m.Az() * wj * vk;
        }
        if (zf(i, j, k, 0, 1, 1)) return 0;
        T Az = m.Az();
        T wt = fy() ? Az * W(i, j, k) : rec_symmetric(m.K, ssch(1, false, i, j, k, Z4), [&](int o) { return Az * W(i, j + o, k); });
        T vr = rec_biased(m.K, bsch(2, false, i, j, k, zb_face(m, k)), bias_left(wt), [&](int o) { return Vv(i, j, k + o); });
        return wt * vr;
    }
    // ---- w momentum fluxes
    T F_Uw(int i, int j, int k) {  // fcf
        if (fx()) return 0;
        const Fld<T> &U = u(), &W = w();
        if (!weno()) {
            T uk = fz() ? U(i, j, k) : T(0.5) * (U(i, j, k - 1) + U(i, j, k));
            T wi = T(0.5) * (W(i - 1, j, k) + W(i, j, k));
            return m.Ax_f(k) * uk * wi;
        }
        if (zf(i, j, k, 1, 0, 1)) return 0;
        T ut = fz() ? m.Ax_c(k) * U(i, j, k)
                    : rec_symmetric(m.K, ssch(2, false, i, j, k, zs_face(m, k)), [&](int o) { return m.Ax_c(k + o) * U(i, j, k + o); });
        T wr = rec_biased(m.K, bsch(0, false, i, j, k, Z5), bias_left(ut), [&](int o) { return W(i + o, j, k); });
        return ut * wr;
    }
    T F_Vw(int i, int j, int k) {  // cff
        if (fy()) return 0;
        const Fld<T> &Vv = v(), &W = w();
        if (!weno()) {
            T vk = fz() ? Vv(i, j, k) : T(0.5) * (Vv(i, j, k - 1) + Vv(i, j, k));
            T wj = T(0.5) * (W(i, j - 1, k) + W(i, j, k));
            return m.Ay_f(k) * vk * wj;
        }
        if (zf(i, j, k, 0, 1, 1)) return 0;
        T vt = fz() ? m.Ay_c(k) * Vv(i, j, k)
                    : rec_symmetric(m.K, ssch(2, false, i, j, k, zs_face(m, k)), [&](int o) { return m.Ay_c(k + o) * Vv(i, j, k + o); });
        T wr = rec_biased(m.K, bsch(1, false, i, j, k, Z5), bias_left(vt), [&](int o) { return W(i, j + o, k); });
        return vt * wr;
    }
    T F_Ww(int i, int j, int k) {  // ccc
        if (fz()) return 0;
        const Fld<T> &W = w();
        if (!weno()) {
            T Wk = T(0.5) * (W(i, j, k) + W(i, j, k + 1));
            return m.Az() * Wk * Wk;
        }
        if (zf(i, j, k, 0, 0, 0)) return 0;
        T Az = m.Az();
        T wt = rec_symmetric(m.K, ssch(2, true, i, j, k, zs_center(m, k)), [&](int o) { return Az * W(i, j, k + 1 + o); });
        T wr = rec_biased(m.K, bsch(2, true, i, j, k, zb_center(m, k)), bias_left(wt), [&](int o) { return W(i, j, k + 1 + o); });
        return wt * wr;
    }
    // ---- tracer fluxes (upwind: advective_tracer_flux_x = Ax·u·c_face; centered: Ax·u·ℑc)
    T F_cx(const Fld<T> &c, int i, int j, int k) {
        if (fx()) return 0;
        T ui = u()(i, j, k);
        if (!weno()) return (m.Ax_c(k) * ui) * (T(0.5) * (c(i - 1, j, k) + c(i, j, k)));
        if (zf(i, j, k, 1, 0, 0)) return 0;
        T cr = rec_biased(m.K, bsch(0, false, i, j, k, Z5), bias_left(ui), [&](int o) { return c(i + o, j, k); });
        return m.Ax_c(k) * ui * cr;
    }
    T F_cy(const Fld<T> &c, int i, int j, int k) {
        if (fy()) return 0;
        T vj = v()(i, j, k);
        if (!weno()) return (m.Ay_c(k) * vj) * (T(0.5) * (c(i, j - 1, k) + c(i, j, k)));
        if (zf(i, j, k, 0, 1, 0)) return 0;
        T cr = rec_biased(m.K, bsch(1, false, i, j, k, Z5), bias_left(vj), [&](int o) { return c(i, j + o, k); });
        return m.Ay_c(k) * vj * cr;
    }
    T F_cz(const Fld<T> &c, int i, int j, int k) {
        if (fz()) return 0;
        T wk = w()(i, j, k);
        if (!weno()) return (m.Az() * wk) * (T(0.5) * (c(i, j, k - 1) + c(i, j, k)));
        if (zf(i, j, k, 0, 0, 1)) return 0;
        T cr = rec_biased(m.K, bsch(2, false, i, j, k, zb_face(m, k)), bias_left(wk), [&](int o) { return c(i, j, k + o); });
        return m.Az() * wk * cr;
    }

    // the differences are conditional δ's on an immersed grid (conditional_differences.jl:15-44)
    T div_Uu(int i, int j, int k) {
        T sx = zd(i, j, k, i - 1, j, k, 0, 0, 0) ? T(0) : F_Uu(i, j, k) - F_Uu(i - 1, j, k);
        T sy = zd(i, j + 1, k, i, j, k, 1, 1, 0) ? T(0) : F_Vu(i, j + 1, k) - F_Vu(i, j, k);
        T sz = zd(i, j, k + 1, i, j, k, 1, 0, 1) ? T(0) : F_Wu(i, j, k + 1) - F_Wu(i, j, k);
        return (T(1) / m.V_c(k)) * ((sx + sy) + sz);
    }
    T div_Uv(int i, int j, int k) {
        T sx = zd(i + 1, j, k, i, j, k, 1, 1, 0) ? T(0) : F_Uv(i + 1, j, k) - F_Uv(i, j, k);
        T sy = zd(i, j, k, i, j - 1, k, 0, 0, 0) ? T(0) : F_Vv(i, j, k) - F_Vv(i, j - 1, k);
        T sz = zd(i, j, k + 1, i, j, k, 0, 1, 1) ? T(0) : F_Wv(i, j, k + 1) - F_Wv(i, j, k);
        return (T(1) / m.V_c(k)) * ((sx + sy) + sz);
    }
    T div_Uw(int i, int j, int k) {
        T sx = zd(i + 1, j, k, i, j, k, 1, 0, 1) ? T(0) : F_Uw(i + 1, j, k) - F_Uw(i, j, k);
        T sy = zd(i, j + 1, k, i, j, k, 0, 1, 1) ? T(0) : F_Vw(i, j + 1, k) - F_Vw(i, j, k);
        T sz = zd(i, j, k, i, j, k - 1, 0, 0, 0) ? T(0) : F_Ww(i, j, k) - F_Ww(i, j, k - 1);
        return (T(1) / m.V_f(k)) * ((sx + sy) + sz);
    }
    T div_Uc(const Fld<T> &c, int i, int j, int k) {
        T sx = zd(i + 1, j, k, i, j, k, 1, 0, 0) ? T(0) : F_cx(c, i + 1, j, k) - F_cx(c, i, j, k);
        T sy = zd(i, j + 1, k, i, j, k, 0, 1, 0) ? T(0) : F_cy(c, i, j + 1, k) - F_cy(c, i, j, k);
        T sz = zd(i, j, k + 1, i, j, k, 0, 0, 1) ? T(0) : F_cz(c, i, j, k + 1) - F_cz(c, i, j, k);
        return (T(1) / m.V_c(k)) * ((sx + sy) + sz);
    }
};

// ----------------------------------------------------------------------------------------
// Closures
// ----------------------------------------------------------------------------------------
template <class T> struct Clo {
    Model<T> &m;
    explicit Clo(Model<T> &mm) : m(mm) {}
    const Fld<T> &u() const { return m.U[0]; }
    const Fld<T> &v() const { return m.U[1]; }
    const Fld<T> &w() const { return m.U[2]; }
    T idx() const { return T(1) / m.g.dx; }
    T idy() const { return T(1) / m.g.dy; }
    bool fz() const { return m.g.topo[2] == F; }
    bool fx() const { return m.g.topo[0] == F; }
    bool fy() const { return m.g.topo[1] == F; }
    // conditional δ (conditional_differences.jl:15-44): zero when either differenced point is an
    // immersed inactive node at the location of the differenced quantity
    bool zd(int i1, int j1, int k1, int i2, int j2, int k2, int lx, int ly, int lz) const {
        if (!m.ib) return false;
        IBP<T> P(m);
        return P.imm_inactive(i1, j1, k1, lx, ly, lz) || P.imm_inactive(i2, j2, k2, lx, ly, lz);
    }
    bool zf(int i, int j, int k, int lx, int ly, int lz) const { return m.ib && IBP<T>(m).imm_peripheral(i, j, k, lx, ly, lz); }
    // derivatives (velocity_tracer_gradients.jl:6-18); Flat directions give zero differences
    T dx_u(int i, int j, int k) const { return fx() || zd(i, j, k, i + 1, j, k, 1, 0, 0) ? T(0) : (u()(i + 1, j, k) - u()(i, j, k)) * idx(); }
    T dy_v(int i, int j, int k) const { return fy() || zd(i, j, k, i, j + 1, k, 0, 1, 0) ? T(0) : (v()(i, j + 1, k) - v()(i, j, k)) * idy(); }
    T dz_w(int i, int j, int k) const { return fz() || zd(i, j, k, i, j, k + 1, 0, 0, 1) ? T(0) : (w()(i, j, k + 1) - w()(i, j, k)) * (T(1) / m.g.dzC(k)); }
    T dx_v(int i, int j, int k) const { return fx() || zd(i, j, k, i - 1, j, k, 0, 1, 0) ? T(0) : (v()(i, j, k) - v()(i - 1, j, k)) * idx(); }
    T dx_w(int i, int j, int k) const { return fx() || zd(i, j, k, i - 1, j, k, 0, 0, 1) ? T(0) : (w()(i, j, k) - w()(i - 1, j, k)) * idx(); }
    T dy_u(int i, int j, int k) const { return fy() || zd(i, j, k, i, j - 1, k, 1, 0, 0) ? T(0) : (u()(i, j, k) - u()(i, j - 1, k)) * idy(); }
    T dy_w(int i, int j, int k) const { return fy() || zd(i, j, k, i, j - 1, k, 0, 0, 1) ? T(0) : (w()(i, j, k) - w()(i, j - 1, k)) * idy(); }
    T dz_u(int i, int j, int k) const { return fz() || zd(i, j, k, i, j, k - 1, 1, 0, 0) ? T(0) : (u()(i, j, k) - u()(i, j, k - 1)) * (T(1) / m.g.dzF(k)); }
    T dz_v(int i, int j, int k) const { return fz() || zd(i, j, k, i, j, k - 1, 0, 1, 0) ? T(0) : (v()(i, j, k) - v()(i, j, k - 1)) * (T(1) / m.g.dzF(k)); }
    T dxc(const Fld<T> &c, int i, int j, int k) const { return fx() || zd(i, j, k, i - 1, j, k, 0, 0, 0) ? T(0) : (c(i, j, k) - c(i - 1, j, k)) * idx(); }
    T dyc(const Fld<T> &c, int i, int j, int k) const { return fy() || zd(i, j, k, i, j - 1, k, 0, 0, 0) ? T(0) : (c(i, j, k) - c(i, j - 1, k)) * idy(); }
    T dzc(const Fld<T> &c, int i, int j, int k) const { return fz() || zd(i, j, k, i, j, k - 1, 0, 0, 0) ? T(0) : (c(i, j, k) - c(i, j, k - 1)) * (T(1) / m.g.dzF(k)); }
    // strain (Σ₁₂ ffc, Σ₁₃ fcf, Σ₂₃ cff)
    T S12(int i, int j, int k) const { return T(0.5) * (dy_u(i, j, k) + dx_v(i, j, k)); }
    T S13(int i, int j, int k) const { return T(0.5) * (dz_u(i, j, k) + dx_w(i, j, k)); }
    T S23(int i, int j, int k) const { return T(0.5) * (dz_v(i, j, k) + dy_w(i, j, k)); }

    // interpolation helpers to ccc from (ff·), (f·f), (·ff) and from faces (Flat → identity)
    template <class Fn> T Ixy_cc(int i, int j, int k, Fn f) const {   // ℑxyᶜᶜᵃ
        auto fxc = [&](int jj) { return fx() ? f(i, jj, k) : T(0.5) * (f(i, jj, k) + f(i + 1, jj, k)); };
        return fy() ? fxc(j) : T(0.5) * (fxc(j) + fxc(j + 1));
    }
    template <class Fn> T Ixz_cc(int i, int j, int k, Fn f) const {   // ℑxzᶜᵃᶜ
        auto fxc = [&](int kk) { return fx() ? f(i, j, kk) : T(0.5) * (f(i, j, kk) + f(i + 1, j, kk)); };
        return fz() ? fxc(k) : T(0.5) * (fxc(k) + fxc(k + 1));
    }
    template <class Fn> T Iyz_cc(int i, int j, int k, Fn f) const {   // ℑyzᵃᶜᶜ
        auto fyc = [&](int kk) { return fy() ? f(i, j, kk) : T(0.5) * (f(i, j, kk) + f(i, j + 1, kk)); };
        return fz() ? fyc(k) : T(0.5) * (fyc(k) + fyc(k + 1));
    }
    template <class Fn> T Ix_c(int i, int j, int k, Fn f) const { return fx() ? f(i, j, k) : T(0.5) * (f(i, j, k) + f(i + 1, j, k)); }
    template <class Fn> T Iy_c(int i, int j, int k, Fn f) const { return fy() ? f(i, j, k) : T(0.5) * (f(i, j, k) + f(i, j + 1, k)); }
    template <class Fn> T Iz_c(int i, int j, int k, Fn f) const { return fz() ? f(i, j, k) : T(0.5) * (f(i, j, k) + f(i, j, k + 1)); }

    // ---------------- AMD (anisotropic_minimum_dissipation.jl:161-358)
    T Dfx() const { return T(2) * m.g.dx; }
    T Dfy() const { return T(2) * m.g.dy; }
    T Dfz(int k) const { return T(2) * m.g.dzC(k); }   // Δᶠz at every location uses Δzᶜᶜᶜ(i, j, k)
    T n_dx_v(int i, int j, int k) const { return Dfx() / Dfy() * dx_v(i, j, k); }
    T n_dy_u(int i, int j, int k) const { return Dfy() / Dfx() * dy_u(i, j, k); }
    T n_dx_w(int i, int j, int k) const { return Dfx() / Dfz(k) * dx_w(i, j, k); }
    T n_dz_u(int i, int j, int k) const { return Dfz(k) / Dfx() * dz_u(i, j, k); }
    T n_dy_w(int i, int j, int k) const { return Dfy() / Dfz(k) * dy_w(i, j, k); }
    T n_dz_v(int i, int j, int k) const { return Dfz(k) / Dfy() * dz_v(i, j, k); }
    T n_dxc(const Fld<T> &c, int i, int j, int k) const { return Dfx() * dxc(c, i, j, k); }
    T n_dyc(const Fld<T> &c, int i, int j, int k) const { return Dfy() * dyc(c, i, j, k); }
    T n_dzc(const Fld<T> &c, int i, int j, int k) const { return Dfz(k) * dzc(c, i, j, k); }
    T n_S12(int i, int j, int k) const { return T(0.5) * (n_dy_u(i, j, k) + n_dx_v(i, j, k)); }
    T n_S13(int i, int j, int k) const { return T(0.5) * (n_dz_u(i, j, k) + n_dx_w(i, j, k)); }
    T n_S23(int i, int j, int k) const { return T(0.5) * (n_dz_v(i, j, k) + n_dy_w(i, j, k)); }

#define L(expr) [&](int a, int b, int c) -> T { return expr; }
    T amd_r(int i, int j, int k) const {
        T S11 = dx_u(i, j, k), S22 = dy_v(i, j, k), S33 = dz_w(i, j, k);
        T t1 = S11 * (dx_u(i, j, k) * dx_u(i, j, k))
             + S22 * (ORC_PLANT == 1 ? Ixz_cc(i, j, k, L(n_dx_v(a, b, c) * n_dx_v(a, b, c)))
                                     : Ixy_cc(i, j, k, L(n_dx_v(a, b, c) * n_dx_v(a, b, c))))
             + S33 * Ixz_cc(i, j, k, L(n_dx_w(a, b, c) * n_dx_w(a, b, c)))
             + T(2) * dx_u(i, j, k) * Ixy_cc(i, j, k, L(n_dx_v(a, b, c) * n_S12(a, b, c)))
             + T(2) * dx_u(i, j, k) * Ixz_cc(i, j, k, L(n_dx_w(a, b, c) * n_S13(a, b, c)))
             + T(2) * Ixy_cc(i, j, k, L(n_dx_v(a, b, c))) * Ixz_cc(i, j, k, L(n_dx_w(a, b, c)))
                    * Iyz_cc(i, j, k, L(n_S23(a, b, c)));
        T t2 = S11 * Ixy_cc(i, j, k, L(n_dy_u(a, b, c) * n_dy_u(a, b, c)))
             + S22 * (dy_v(i, j, k) * dy_v(i, j, k))
             + S33 * Iyz_cc(i, j, k, L(n_dy_w(a, b, c) * n_dy_w(a, b, c)))
             + T(2) * dy_v(i, j, k) * Ixy_cc(i, j, k, L(n_dy_u(a, b, c) * n_S12(a, b, c)))
             + T(2) * Ixy_cc(i, j, k, L(n_dy_u(a, b, c))) * Iyz_cc(i, j, k, L(n_dy_w(a, b, c)))
                    * Ixz_cc(i, j, k, L(n_S13(a, b, c)))
             + T(2) * dy_v(i, j, k) * Iyz_cc(i, j, k, L(n_dy_w(a, b, c) * n_S23(a, b, c)));
        T t3 = S11 * Ixz_cc(i, j, k, L(n_dz_u(a, b, c) * n_dz_u(a, b, c)))
             + S22 * Iyz_cc(i, j, k, L(n_dz_v(a, b, c) * n_dz_v(a, b, c)))
             + S33 * (dz_w(i, j, k) * dz_w(i, j, k))
             + T(2) * Ixz_cc(i, j, k, L(n_dz_u(a, b, c))) * Iyz_cc(i, j, k, L(n_dz_v(a, b, c)))
                    * Ixy_cc(i, j, k, L(n_S12(a, b, c)))
             + T(2) * dz_w(i, j, k) * Ixz_cc(i, j, k, L(n_dz_u(a, b, c) * n_S13(a, b, c)))
             + T(2) * dz_w(i, j, k) * Iyz_cc(i, j, k, L(n_dz_v(a, b, c) * n_S23(a, b, c)));
        return t1 + t2 + t3;
    }
    T amd_q(int i, int j, int k) const {
        return dx_u(i, j, k) * dx_u(i, j, k) + dy_v(i, j, k) * dy_v(i, j, k) + dz_w(i, j, k) * dz_w(i, j, k)
             + Ixy_cc(i, j, k, L(n_dx_v(a, b, c) * n_dx_v(a, b, c)))
             + Ixy_cc(i, j, k, L(n_dy_u(a, b, c) * n_dy_u(a, b, c)))
             + Ixz_cc(i, j, k, L(n_dx_w(a, b, c) * n_dx_w(a, b, c)))
             + Ixz_cc(i, j, k, L(n_dz_u(a, b, c) * n_dz_u(a, b, c)))
             + Iyz_cc(i, j, k, L(n_dy_w(a, b, c) * n_dy_w(a, b, c)))
             + Iyz_cc(i, j, k, L(n_dz_v(a, b, c) * n_dz_v(a, b, c)));
    }
    T amd_delta2(int k) const {
        T a = T(1) / (Dfx() * Dfx()), b = T(1) / (Dfy() * Dfy()), c = T(1) / (Dfz(k) * Dfz(k));
        return T(3) / ((a + b) + c);
    }
    T amd_nu(int i, int j, int k) const {
        T q = amd_q(i, j, k);
        if (q == T(0)) return T(0);
        T r = amd_r(i, j, k);
        T Cb_zeta = T(0);   // Cb = nothing
        T nu = -(T)m.cfg.closure_params[0] * amd_delta2(k) * (r - Cb_zeta) / q;
        return std::max(T(0), nu);
    }
    T amd_kappa(const Fld<T> &cf, int i, int j, int k) const {
        auto ncx = [&](int a, int b, int c) { return n_dxc(cf, a, b, c); };
        auto ncy = [&](int a, int b, int c) { return n_dyc(cf, a, b, c); };
        auto ncz = [&](int a, int b, int c) { return n_dzc(cf, a, b, c); };
        auto ncx2 = [&](int a, int b, int c) { T d = n_dxc(cf, a, b, c); return d * d; };
        auto ncy2 = [&](int a, int b, int c) { T d = n_dyc(cf, a, b, c); return d * d; };
        auto ncz2 = [&](int a, int b, int c) { T d = n_dzc(cf, a, b, c); return d * d; };
        T s = Ix_c(i, j, k, ncx2) + Iy_c(i, j, k, ncy2) + Iz_c(i, j, k, ncz2);
        if (s == T(0)) return T(0);
        T cx_ux = dx_u(i, j, k) * Ix_c(i, j, k, ncx2)
                + Ixy_cc(i, j, k, L(n_dx_v(a, b, c))) * Ix_c(i, j, k, ncx) * Iy_c(i, j, k, ncy)
                + Ixz_cc(i, j, k, L(n_dx_w(a, b, c))) * Ix_c(i, j, k, ncx) * Iz_c(i, j, k, ncz);
        T cy_uy = Ixy_cc(i, j, k, L(n_dy_u(a, b, c))) * Iy_c(i, j, k, ncy) * Ix_c(i, j, k, ncx)
                + dy_v(i, j, k) * Iy_c(i, j, k, ncy2)
                // the reference interpolates ∂y_w (a cff quantity) with ℑxzᶜᵃᶜ here (anisotropic_minimum_dissipation.jl:343)
                + Ixz_cc(i, j, k, L(n_dy_w(a, b, c))) * Iy_c(i, j, k, ncy) * Iz_c(i, j, k, ncz);
        T cz_uz = Ixz_cc(i, j, k, L(n_dz_u(a, b, c))) * Iz_c(i, j, k, ncz) * Ix_c(i, j, k, ncx)
                + Iyz_cc(i, j, k, L(n_dz_v(a, b, c))) * Iz_c(i, j, k, ncz) * Iy_c(i, j, k, ncy)
                + dz_w(i, j, k) * Iz_c(i, j, k, ncz2);
        T theta = cx_ux + cy_uy + cz_uz;
        T kappa = -(T)m.cfg.closure_params[1] * amd_delta2(k) * theta / s;
        return std::max(T(0), kappa);
    }
#undef L

    // ---------------- Smagorinsky-Lilly (smagorinsky.jl:90-104, lilly_coefficient.jl:129-142, scale_invariant_operators.jl:10-13)
    T dzb(int i, int j, int k) const;   // ∂z_b at ccf
    T smag_nu(int i, int j, int k) const {
        T tr = dx_u(i, j, k) * dx_u(i, j, k) + dy_v(i, j, k) * dy_v(i, j, k) + dz_w(i, j, k) * dz_w(i, j, k);
        T S2 = tr + T(2) * Ixy_cc(i, j, k, [&](int a, int b, int c) { T s = S12(a, b, c); return s * s; })
                  + T(2) * Ixz_cc(i, j, k, [&](int a, int b, int c) { T s = S13(a, b, c); return s * s; })
                  + T(2) * Iyz_cc(i, j, k, [&](int a, int b, int c) { T s = S23(a, b, c); return s * s; });
        T D3 = (m.g.dx * m.g.dy) * m.g.dzC(k);
        T Df = std::cbrt(D3);
        T N2 = Iz_c(i, j, k, [&](int a, int b, int c) { return dzb(a, b, c); });
        T c0 = (T)m.cfg.closure_params[0], cb = (T)m.cfg.closure_params[1];
        T N2p = std::max(T(0), N2);
        T s2 = T(1) - std::min(T(1), cb * N2p / S2);
        T sig = (S2 == T(0)) ? T(0) : std::sqrt(s2);
        T cs2 = sig * (c0 * c0);
        return cs2 * (Df * Df) * std::sqrt(T(2) * S2);
    }

    // ---------------- viscous fluxes (abstract_scalar_diffusivity_closure.jl:214-262, 343-353)
    T nu_ccc(int i, int j, int k) const;
    T nu_ffc(int i, int j, int k) const;
    T nu_fcf(int i, int j, int k) const;
    T nu_cff(int i, int j, int k) const;
    T kap_fcc(int n, int i, int j, int k) const;
    T kap_cfc(int n, int i, int j, int k) const;
    T kap_ccf(int n, int i, int j, int k) const;

    // viscous fluxes; zero on the immersed periphery of their location (immersed_diffusive_fluxes.jl:25-42)
    T f_ux(int i, int j, int k) const { return zf(i, j, k, 0, 0, 0) ? T(0) : T(-2) * (nu_ccc(i, j, k) * dx_u(i, j, k)); }
    T f_uy(int i, int j, int k) const { return zf(i, j, k, 1, 1, 0) ? T(0) : T(-2) * (nu_ffc(i, j, k) * S12(i, j, k)); }
    T f_uz(int i, int j, int k) const { return zf(i, j, k, 1, 0, 1) ? T(0) : T(-2) * (nu_fcf(i, j, k) * S13(i, j, k)); }
    T f_vx(int i, int j, int k) const { return zf(i, j, k, 1, 1, 0) ? T(0) : T(-2) * (nu_ffc(i, j, k) * S12(i, j, k)); }
    T f_vy(int i, int j, int k) const { return zf(i, j, k, 0, 0, 0) ? T(0) : T(-2) * (nu_ccc(i, j, k) * dy_v(i, j, k)); }
    T f_vz(int i, int j, int k) const { return zf(i, j, k, 0, 1, 1) ? T(0) : T(-2) * (nu_cff(i, j, k) * S23(i, j, k)); }
    T f_wx(int i, int j, int k) const { return zf(i, j, k, 1, 0, 1) ? T(0) : T(-2) * (nu_fcf(i, j, k) * S13(i, j, k)); }
    T f_wy(int i, int j, int k) const { return zf(i, j, k, 0, 1, 1) ? T(0) : T(-2) * (nu_cff(i, j, k) * S23(i, j, k)); }
    T f_wz(int i, int j, int k) const { return zf(i, j, k, 0, 0, 0) ? T(0) : T(-2) * (nu_ccc(i, j, k) * dz_w(i, j, k)); }

    // closure_kernel_operators.jl:20-46
    T div_tau1(int i, int j, int k) const {
        T sx = fx() || zd(i, j, k, i - 1, j, k, 0, 0, 0) ? T(0) : m.Ax_c(k) * f_ux(i, j, k) - m.Ax_c(k) * f_ux(i - 1, j, k);
        T sy = fy() || zd(i, j + 1, k, i, j, k, 1, 1, 0) ? T(0) : m.Ay_c(k) * f_uy(i, j + 1, k) - m.Ay_c(k) * f_uy(i, j, k);
        T sz = fz() || zd(i, j, k + 1, i, j, k, 1, 0, 1) ? T(0) : m.Az() * f_uz(i, j, k + 1) - m.Az() * f_uz(i, j, k);
        return (T(1) / m.V_c(k)) * ((sx + sy) + sz);
    }
    T div_tau2(int i, int j, int k) const {
        T sx = fx() || zd(i + 1, j, k, i, j, k, 1, 1, 0) ? T(0) : m.Ax_c(k) * f_vx(i + 1, j, k) - m.Ax_c(k) * f_vx(i, j, k);
        T sy = fy() || zd(i, j, k, i, j - 1, k, 0, 0, 0) ? T(0) : m.Ay_c(k) * f_vy(i, j, k) - m.Ay_c(k) * f_vy(i, j - 1, k);
        T sz = fz() || zd(i, j, k + 1, i, j, k, 0, 1, 1) ? T(0) : m.Az() * f_vz(i, j, k + 1) - m.Az() * f_vz(i, j, k);
        return (T(1) / m.V_c(k)) * ((sx + sy) + sz);
    }
    T div_tau3(int i, int j, int k) const {
        T sx = fx() || zd(i + 1, j, k, i, j, k, 1, 0, 1) ? T(0) : m.Ax_f(k) * f_wx(i + 1, j, k) - m.Ax_f(k) * f_wx(i, j, k);
        T sy = fy() || zd(i, j + 1, k, i, j, k, 0, 1, 1) ? T(0) : m.Ay_f(k) * f_wy(i, j + 1, k) - m.Ay_f(k) * f_wy(i, j, k);
        T sz = fz() || zd(i, j, k, i, j, k - 1, 0, 0, 0) ? T(0) : m.Az() * f_wz(i, j, k) - m.Az() * f_wz(i, j, k - 1);
        return (T(1) / m.V_f(k)) * ((sx + sy) + sz);
    }
    T div_q(int n, const Fld<T> &c, int i, int j, int k) const {
        auto qx = [&](int a) { return zf(a, j, k, 1, 0, 0) ? T(0) : -kap_fcc(n, a, j, k) * dxc(c, a, j, k); };
        auto qy = [&](int b) { return zf(i, b, k, 0, 1, 0) ? T(0) : -kap_cfc(n, i, b, k) * dyc(c, i, b, k); };
        auto qz = [&](int cc) { return zf(i, j, cc, 0, 0, 1) ? T(0) : -kap_ccf(n, i, j, cc) * dzc(c, i, j, cc); };
        T sx = fx() || zd(i + 1, j, k, i, j, k, 1, 0, 0) ? T(0) : m.Ax_c(k) * qx(i + 1) - m.Ax_c(k) * qx(i);
        T sy = fy() || zd(i, j + 1, k, i, j, k, 0, 1, 0) ? T(0) : m.Ay_c(k) * qy(j + 1) - m.Ay_c(k) * qy(j);
        T sz = fz() || zd(i, j, k + 1, i, j, k, 0, 0, 1) ? T(0) : m.Az() * qz(k + 1) - m.Az() * qz(k);
        return (T(1) / m.V_c(k)) * ((sx + sy) + sz);
    }
};

template <class T> static T buoyancy_pert(const Model<T> &m, int i, int j, int k) {
    if (m.cfg.buoyancy == OM_BUOYANCY_TRACER) return m.U[3 + m.cfg.buoyancy_tracer](i, j, k);
    if (m.cfg.buoyancy == OM_BUOYANCY_LINEAR_EOS) {
        T g = (T)m.cfg.gravitational_acceleration, a = (T)m.cfg.thermal_expansion, b = (T)m.cfg.haline_contraction;
        return g * (a * m.U[3 + m.cfg.T_tracer](i, j, k) - b * m.U[3 + m.cfg.S_tracer](i, j, k));
    }
    return T(0);
}

template <class T> T Clo<T>::dzb(int i, int j, int k) const {
    if (fz()) return T(0);
    if (m.cfg.buoyancy == OM_BUOYANCY_TRACER) return dzc(m.U[3 + m.cfg.buoyancy_tracer], i, j, k);
    if (m.cfg.buoyancy == OM_BUOYANCY_LINEAR_EOS) {
        T g = (T)m.cfg.gravitational_acceleration, a = (T)m.cfg.thermal_expansion, b = (T)m.cfg.haline_contraction;
        return g * (a * dzc(m.U[3 + m.cfg.T_tracer], i, j, k) - b * dzc(m.U[3 + m.cfg.S_tracer], i, j, k));
    }
    return T(0);
}

template <class T> T Clo<T>::nu_ccc(int i, int j, int k) const {
    if (m.cfg.closure == OM_CLOSURE_SCALAR_DIFFUSIVITY) return (T)m.cfg.closure_params[0];
    return m.nu(i, j, k);
}
template <class T> T Clo<T>::nu_ffc(int i, int j, int k) const {
    if (m.cfg.closure == OM_CLOSURE_SCALAR_DIFFUSIVITY) return (T)m.cfg.closure_params[0];
    auto fxf = [&](int b) { return fx() ? m.nu(i, b, k) : T(0.5) * (m.nu(i - 1, b, k) + m.nu(i, b, k)); };
    return fy() ? fxf(j) : T(0.5) * (fxf(j - 1) + fxf(j));
}
template <class T> T Clo<T>::nu_fcf(int i, int j, int k) const {
    if (m.cfg.closure == OM_CLOSURE_SCALAR_DIFFUSIVITY) return (T)m.cfg.closure_params[0];
    auto fxf = [&](int c) { return fx() ? m.nu(i, j, c) : T(0.5) * (m.nu(i - 1, j, c) + m.nu(i, j, c)); };
    return fz() ? fxf(k) : T(0.5) * (fxf(k - 1) + fxf(k));
}
template <class T> T Clo<T>::nu_cff(int i, int j, int k) const {
    if (m.cfg.closure == OM_CLOSURE_SCALAR_DIFFUSIVITY) return (T)m.cfg.closure_params[0];
    auto fyf = [&](int c) { return fy() ? m.nu(i, j, c) : T(0.5) * (m.nu(i, j - 1, c) + m.nu(i, j, c)); };
    return fz() ? fyf(k) : T(0.5) * (fyf(k - 1) + fyf(k));
}
// tracer diffusivities: AMD κₑ per tracer; Smagorinsky ℑ(νₑ)/Pr; constant κ
template <class T> T Clo<T>::kap_fcc(int n, int i, int j, int k) const {
    if (m.cfg.closure == OM_CLOSURE_SCALAR_DIFFUSIVITY) return (T)m.cfg.closure_params[1];
    if (m.cfg.closure == OM_CLOSURE_AMD) return fx() ? m.kap[n](i, j, k) : T(0.5) * (m.kap[n](i - 1, j, k) + m.kap[n](i, j, k));
    T nf = fx() ? m.nu(i, j, k) : T(0.5) * (m.nu(i - 1, j, k) + m.nu(i, j, k));
    return nf / (T)m.cfg.closure_params[2];
}
template <class T> T Clo<T>::kap_cfc(int n, int i, int j, int k) const {
    if (m.cfg.closure == OM_CLOSURE_SCALAR_DIFFUSIVITY) return (T)m.cfg.closure_params[1];
    if (m.cfg.closure == OM_CLOSURE_AMD) return fy() ? m.kap[n](i, j, k) : T(0.5) * (m.kap[n](i, j - 1, k) + m.kap[n](i, j, k));
    T nf = fy() ? m.nu(i, j, k) : T(0.5) * (m.nu(i, j - 1, k) + m.nu(i, j, k));
    return nf / (T)m.cfg.closure_params[2];
}
template <class T> T Clo<T>::kap_ccf(int n, int i, int j, int k) const {
    if (m.cfg.closure == OM_CLOSURE_SCALAR_DIFFUSIVITY) return (T)m.cfg.closure_params[1];
    if (m.cfg.closure == OM_CLOSURE_AMD) return fz() ? m.kap[n](i, j, k) : T(0.5) * (m.kap[n](i, j, k - 1) + m.kap[n](i, j, k));
    T nf = fz() ? m.nu(i, j, k) : T(0.5) * (m.nu(i, j, k - 1) + m.nu(i, j, k));
    return nf / (T)m.cfg.closure_params[2];
}

// ----------------------------------------------------------------------------------------
// Model phases
// ----------------------------------------------------------------------------------------
template <class T> static void fill_all_prognostic_halos(Model<T> &m, bool fill_normal) {
    for (int f = 0; f < 3 + m.nt; ++f) fill_prognostic_halos(m, f, fill_normal);
}

// compute_closure_fields! + fill_halo_regions!(closure_fields)
template <class T> static void compute_closure_fields(Model<T> &m) {
    if (m.cfg.closure != OM_CLOSURE_AMD && m.cfg.closure != OM_CLOSURE_SMAGORINSKY_LILLY) return;
    Clo<T> C(m);
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
        if (m.cfg.closure == OM_CLOSURE_AMD) {
            m.nu(i, j, k) = C.amd_nu(i, j, k);
            for (int n = 0; n < m.nt; ++n) m.kap[n](i, j, k) = C.amd_kappa(m.U[3 + n], i, j, k);
        } else {
            m.nu(i, j, k) = C.smag_nu(i, j, k);
        }
    }
    fill_aux_halos(m, m.nu);
    if (m.cfg.closure == OM_CLOSURE_AMD) for (int n = 0; n < m.nt; ++n) fill_aux_halos(m, m.kap[n]);
}

// update_hydrostatic_pressure.jl:11-21 (columns over the interior; halos filled afterwards)
template <class T> static void update_hydrostatic_pressure(Model<T> &m) {
    if (!m.has_pHY) return;
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
    auto Izb = [&](int i, int j, int k) { return T(0.5) * (buoyancy_pert(m, i, j, k - 1) + buoyancy_pert(m, i, j, k)); };
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
        T p = -Izb(i, j, Nz + 1) * m.g.dzF(ORC_PLANT == 2 ? Nz : Nz + 1);
        m.pHY(i, j, Nz) = p;
        for (int k = Nz - 1; k >= 1; --k) {
            p -= Izb(i, j, k + 1) * m.g.dzF(k + 1);
            m.pHY(i, j, k) = p;
        }
    }
    fill_aux_halos(m, m.pHY);
}

// nonhydrostatic_tendency_kernel_functions.jl:71-302
template <class T> static void compute_tendencies(Model<T> &m) {
    Adv<T> A(m);
    Clo<T> C(m);
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
    const bool clo = m.cfg.closure != OM_CLOSURE_NONE;
    const bool adv = m.cfg.advection != OM_ADV_NONE;
    const bool cor = m.cfg.coriolis == 1;
    const T f = (T)m.cfg.f;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
        // ---- u
        {
            T G = 0;
            if (adv) G -= A.div_Uu(i, j, k);
            if (cor) {
                // - x_f_cross_U = + f ℑxyᶠᶜᵃ(Ay v) / Ayᶠᶜᶜ   (coriolis_schemes.jl:67)
                // masked_Ay_qᶜᶠᶜ: zero on the immersed periphery (coriolis_schemes.jl:56-59)
                auto q = [&](int a, int b) { return IBP<T>(m).imm_peripheral(a, b, k, 0, 1, 0) ? T(0) : m.Ay_c(k) * m.U[1](a, b, k); };
                T vv = m.g.topo[0] == F ? q(i, j) + q(i, j + 1) : T(0);
                T Ixf_j = m.g.topo[0] == F ? q(i, j) : T(0.5) * (q(i - 1, j) + q(i, j));
                T Ixf_j1 = m.g.topo[0] == F ? q(i, j + 1) : T(0.5) * (q(i - 1, j + 1) + q(i, j + 1));
                (void)vv;
                T Ixy = m.g.topo[1] == F ? Ixf_j : T(0.5) * (Ixf_j + Ixf_j1);
                T fI = T(0.5) * (f + f);
                G -= -fI * Ixy * (T(1) / m.Ay_c(k));
            }
            if (m.has_pHY) G -= (C.zd(i, j, k, i - 1, j, k, 0, 0, 0) ? T(0) : m.pHY(i, j, k) - m.pHY(i - 1, j, k)) * (T(1) / m.g.dx);
            if (clo) G -= C.div_tau1(i, j, k);
            m.Gn[0](i, j, k) = G;
        }
        // ---- v
        {
            T G = 0;
            if (adv) G -= A.div_Uv(i, j, k);
            if (cor) {
                // - y_f_cross_U = - f ℑxyᶜᶠᵃ(Ax u) / Axᶜᶠᶜ   (coriolis_schemes.jl:68)
                auto q = [&](int a, int b) { return IBP<T>(m).imm_peripheral(a, b, k, 1, 0, 0) ? T(0) : m.Ax_c(k) * m.U[0](a, b, k); };
                T Ixc_jm = m.g.topo[0] == F ? q(i, j - 1) : T(0.5) * (q(i, j - 1) + q(i + 1, j - 1));
                T Ixc_j = m.g.topo[0] == F ? q(i, j) : T(0.5) * (q(i, j) + q(i + 1, j));
                T Ixy = m.g.topo[1] == F ? Ixc_j : T(0.5) * (Ixc_jm + Ixc_j);
                T fI = T(0.5) * (f + f);
                G -= fI * Ixy * (T(1) / m.Ax_c(k));
            }
            if (m.has_pHY) G -= (C.zd(i, j, k, i, j - 1, k, 0, 0, 0) ? T(0) : m.pHY(i, j, k) - m.pHY(i, j - 1, k)) * (T(1) / m.g.dy);
            if (clo) G -= C.div_tau2(i, j, k);
            m.Gn[1](i, j, k) = G;
        }
        // ---- w
        {
            T G = 0;
            if (adv) G -= A.div_Uw(i, j, k);
            // buoyancy enters w only when there is no hydrostatic pressure anomaly
            if (!m.has_pHY && m.cfg.buoyancy != OM_BUOYANCY_NONE && !m.flatz())
                G += T(0.5) * (buoyancy_pert(m, i, j, k - 1) + buoyancy_pert(m, i, j, k));
            if (clo) G -= C.div_tau3(i, j, k);
            m.Gn[2](i, j, k) = G;
        }
        // ---- tracers
        for (int n = 0; n < m.nt; ++n) {
            T G = 0;
            if (adv) G -= A.div_Uc(m.U[3 + n], i, j, k);
            if (clo) G -= C.div_q(n, m.U[3 + n], i, j, k);
            m.Gn[3 + n](i, j, k) = G;
        }
    }
}

// compute_flux_bcs.jl:113-162 (z only; x, y are Periodic or Flat)
template <class T> static void compute_flux_bc_tendencies(Model<T> &m) {
    if (!m.bz()) return;
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
    for (int f = 0; f < 3 + m.nt; ++f) {
        if (f == 2) continue;
        bool bot = m.cfg.bc_bottom[f] == OM_BC_FLUX, top = m.cfg.bc_top[f] == OM_BC_FLUX;
        if (!bot && !top) continue;
        T qb = (T)m.cfg.bc_bottom_value[f], qt = (T)m.cfg.bc_top_value[f];
        for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
            if (bot) m.Gn[f](i, j, 1) += qb * m.Az() / m.V_c(1);
            if (top) m.Gn[f](i, j, Nz) -= qt * m.Az() / m.V_c(Nz);
        }
    }
}

// ----------------------------------------------------------------------------------------
// FFT machinery (transforms along one axis of a complex 3D array, x fastest)
// ----------------------------------------------------------------------------------------
typedef std::complex<double> cplx;

static void fft_inplace(std::vector<cplx> &a, bool inverse) {
    const size_t n = a.size();
    if (n <= 1) return;
    if ((n & (n - 1)) == 0) {
        for (size_t i = 1, j = 0; i < n; ++i) {
            size_t bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) std::swap(a[i], a[j]);
        }
        for (size_t len = 2; len <= n; len <<= 1) {
            double ang = 2 * M_PI / (double)len * (inverse ? 1 : -1);
            for (size_t i = 0; i < n; i += len)
                for (size_t q = 0; q < len / 2; ++q) {
                    cplx w(std::cos(ang * q), std::sin(ang * q));
                    cplx x = a[i + q], y = a[i + q + len / 2] * w;
                    a[i + q] = x + y; a[i + q + len / 2] = x - y;
                }
        }
    } else {
        std::vector<cplx> out(n);
        for (size_t k = 0; k < n; ++k) {
            cplx s = 0;
            for (size_t t = 0; t < n; ++t) {
                double ang = 2 * M_PI * (double)((k * t) % n) / (double)n * (inverse ? 1 : -1);
                s += a[t] * cplx(std::cos(ang), std::sin(ang));
            }
            out[k] = s;
        }
        a.swap(out);
    }
    if (inverse) for (auto &x : a) x /= (double)n;
}

// DCT-II (FFTW REDFT10: X_k = 2 Σ x_n cos(πk(2n+1)/2N)) and its inverse (REDFT01 / 2N) via a 2N FFT
static void dct2_inplace(std::vector<cplx> &x) {
    const size_t N = x.size();
    std::vector<cplx> e(2 * N);
    for (size_t n = 0; n < N; ++n) { e[n] = x[n]; e[2 * N - 1 - n] = x[n]; }
    fft_inplace(e, false);
    for (size_t k = 0; k < N; ++k) x[k] = e[k] * std::polar(1.0, -M_PI * (double)k / (2.0 * N));
}
static void idct2_inplace(std::vector<cplx> &X) {
    const size_t N = X.size();
    std::vector<cplx> Y(2 * N);
    for (size_t k = 0; k < N; ++k) Y[k] = X[k] * std::polar(1.0, M_PI * (double)k / (2.0 * N));
    Y[N] = 0;
    for (size_t k = 1; k < N; ++k) Y[2 * N - k] = X[k] * std::polar(1.0, -M_PI * (double)k / (2.0 * N));
    fft_inplace(Y, true);
    for (size_t n = 0; n < N; ++n) X[n] = Y[n];
}

// Apply a 1D transform along `dim` to a complex array of shape (Nx, Ny, Nz) (x fastest).
template <class Fn> static void along_axis(std::vector<cplx> &a, const int N[3], int dim, Fn fn) {
    const int n0 = N[0], n1 = N[1], n2 = N[2];
    const int len = N[dim];
    const int other1 = dim == 0 ? n1 : n0, other2 = dim == 2 ? n1 : n2;
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < other2; ++b) for (int a0 = 0; a0 < other1; ++a0) {
        std::vector<cplx> line(len);
        auto index = [&](int t) -> size_t {
            int i, j, k;
            if (dim == 0) { i = t; j = a0; k = b; }
            else if (dim == 1) { i = a0; j = t; k = b; }
            else { i = a0; j = b; k = t; }
            return (size_t)i + (size_t)n0 * ((size_t)j + (size_t)n1 * k);
        };
        for (int t = 0; t < len; ++t) line[t] = a[index(t)];
        fn(line);
        for (int t = 0; t < len; ++t) a[index(t)] = line[t];
    }
}

// poisson_eigenvalues.jl:8-32, computed in double and rounded to FT
template <class T> static std::vector<T> eigenvalues(int N, double L, int topo) {
    std::vector<T> lam(N, T(0));
    for (int i = 1; i <= N; ++i) {
        double v = 0;
        if (topo == P) v = std::pow(2 * std::sin((i - 1) * M_PI / N) / (L / N), 2);
        else if (topo == B) v = std::pow(2 * std::sin((i - 1) * M_PI / (2.0 * N)) / (L / N), 2);
        lam[i - 1] = (T)v;
    }
    return lam;
}

// solve!(ϕ, ::FFTBasedPoissonSolver, b): forward transforms, ϕ = -b/(λx+λy+λz), ϕ[1,1,1] = 0, backward
template <class T> static void fft_poisson(const Model<T> &m, std::vector<cplx> &a) {
    const int *N = m.g.N;
    // Bounded dimensions are transformed first in the forward direction; the result is
    // independent of the order (the transforms along different axes commute).
    for (int d = 0; d < 3; ++d) {
        if (m.g.topo[d] == P) along_axis(a, N, d, [](std::vector<cplx> &l) { fft_inplace(l, false); });
        else if (m.g.topo[d] == B) along_axis(a, N, d, [](std::vector<cplx> &l) { dct2_inplace(l); });
    }
    auto lx = eigenvalues<T>(N[0], (double)m.g.Lx, m.g.topo[0]);
    auto ly = eigenvalues<T>(N[1], (double)m.g.Ly, m.g.topo[1]);
    auto lz = eigenvalues<T>(N[2], (double)m.g.Lz, m.g.topo[2]);
    for (int k = 0; k < N[2]; ++k) for (int j = 0; j < N[1]; ++j) for (int i = 0; i < N[0]; ++i) {
        size_t id = (size_t)i + (size_t)N[0] * ((size_t)j + (size_t)N[1] * k);
        T den = (lx[i] + ly[j]) + lz[k];
        a[id] = -a[id] / (double)den;
    }
    a[0] = 0;
    for (int d = 2; d >= 0; --d) {
        if (m.g.topo[d] == P) along_axis(a, N, d, [](std::vector<cplx> &l) { fft_inplace(l, true); });
        else if (m.g.topo[d] == B) along_axis(a, N, d, [](std::vector<cplx> &l) { idct2_inplace(l); });
    }
}

// FourierTridiagonalPoissonSolver (fourier_tridiagonal_poisson_solver.jl:87-290) with the
// tridiagonal direction z: rhs·Δzᶜ, x/y transforms, Thomas in z, backward, ϕ -= mean(ϕ).
template <class T> static void fourier_tridiagonal_poisson(const Model<T> &m, std::vector<cplx> &a) {
    const int *N = m.g.N;
    for (int d = 0; d < 2; ++d) {
        if (m.g.topo[d] == P) along_axis(a, N, d, [](std::vector<cplx> &l) { fft_inplace(l, false); });
        else if (m.g.topo[d] == B) along_axis(a, N, d, [](std::vector<cplx> &l) { dct2_inplace(l); });
    }
    auto lx = eigenvalues<T>(N[0], (double)m.g.Lx, m.g.topo[0]);
    auto ly = eigenvalues<T>(N[1], (double)m.g.Ly, m.g.topo[1]);
    const int Nz = N[2];
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < N[1]; ++j) for (int i = 0; i < N[0]; ++i) {
        // batched_tridiagonal_solver.jl:211-243 (Thomas); diagonals from :199-229
        std::vector<double> lower(Nz), diag(Nz), upper(Nz), tt(Nz);
        std::vector<cplx> f(Nz), phi(Nz);
        for (int k = 1; k <= Nz; ++k) {
            T dzc = m.g.dzC(k);
            T D;
            if (k == 1) D = -(T(1) / m.g.dzF(2)) - dzc * (lx[i] + ly[j]);
            else if (k == Nz) D = -(T(1) / m.g.dzF(Nz)) - dzc * (lx[i] + ly[j]);
            else D = -((T(1) / m.g.dzF(k + 1)) + (T(1) / m.g.dzF(k))) - dzc * (lx[i] + ly[j]);
            diag[k - 1] = (double)D;
            if (k < Nz) { lower[k - 1] = (double)(T(1) / m.g.dzF(k + 1)); upper[k - 1] = lower[k - 1]; }
            f[k - 1] = a[(size_t)i + (size_t)N[0] * ((size_t)j + (size_t)N[1] * (k - 1))];
        }
        // the (0,0) mode is singular: its last pivot is a rounding residue and the reference
        // elides that update (keeping a stale value); the constant it adds is removed with the mean.
        double beta = diag[0];
        phi[0] = f[0] / beta;
        for (int k = 1; k < Nz; ++k) {
            tt[k] = upper[k - 1] / beta;
            beta = diag[k] - lower[k - 1] * tt[k];
            if (std::abs(beta) > 10 * std::numeric_limits<T>::epsilon()) phi[k] = (f[k] - lower[k - 1] * phi[k - 1]) / beta;
            else phi[k] = 0;
        }
        for (int k = Nz - 2; k >= 0; --k) phi[k] = phi[k] - tt[k + 1] * phi[k + 1];
        for (int k = 1; k <= Nz; ++k) a[(size_t)i + (size_t)N[0] * ((size_t)j + (size_t)N[1] * (k - 1))] = phi[k - 1];
    }
    for (int d = 1; d >= 0; --d) {
        if (m.g.topo[d] == P) along_axis(a, N, d, [](std::vector<cplx> &l) { fft_inplace(l, true); });
        else if (m.g.topo[d] == B) along_axis(a, N, d, [](std::vector<cplx> &l) { idct2_inplace(l); });
    }
    cplx mean = 0;
    for (auto &x : a) mean += x;
    mean /= (double)a.size();
    for (auto &x : a) x -= mean;
}

// rhs = divᶜᶜᶜ(U) (solve_for_pressure.jl:12-18 / :36-42 scaled by Δzᶜ for Fourier-tridiagonal)
template <class T> static T divergence(const Model<T> &m, int i, int j, int k) {
    IBP<T> P(m);
    auto zd = [&](int i1, int j1, int k1, int i2, int j2, int k2, int lx, int ly, int lz) {
        return m.ib && (P.imm_inactive(i1, j1, k1, lx, ly, lz) || P.imm_inactive(i2, j2, k2, lx, ly, lz));
    };
    T dxs = m.g.topo[0] == F || zd(i, j, k, i + 1, j, k, 1, 0, 0) ? T(0) : m.Ax_c(k) * m.U[0](i + 1, j, k) - m.Ax_c(k) * m.U[0](i, j, k);
    T dys = m.g.topo[1] == F || zd(i, j, k, i, j + 1, k, 0, 1, 0) ? T(0) : m.Ay_c(k) * m.U[1](i, j + 1, k) - m.Ay_c(k) * m.U[1](i, j, k);
    T dzs = m.g.topo[2] == F || zd(i, j, k, i, j, k + 1, 0, 0, 1) ? T(0) : m.Az() * m.U[2](i, j, k + 1) - m.Az() * m.U[2](i, j, k);
    return (T(1) / m.V_c(k)) * ((dxs + dys) + dzs);
}

template <class T> static void cg_solve(Model<T> &m, Fld<T> &x, Fld<T> &b);

template <class T> static void solve_pressure_field(Model<T> &m, const std::vector<T> &rhs, Fld<T> &phi) {
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
    if (m.cfg.pressure_solver == OM_SOLVER_CONJUGATE_GRADIENT) {
        for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
            m.cg_b(i, j, k) = rhs[(size_t)(i - 1) + (size_t)Nx * ((size_t)(j - 1) + (size_t)Ny * (k - 1))] * m.V_c(k);
            phi(i, j, k) = T(0);
        }
        cg_solve(m, phi, m.cg_b);
        return;
    }
    std::vector<cplx> a((size_t)Nx * Ny * Nz);
    const bool ft = m.cfg.pressure_solver == OM_SOLVER_FOURIER_TRIDIAGONAL;
    for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
        size_t id = (size_t)(i - 1) + (size_t)Nx * ((size_t)(j - 1) + (size_t)Ny * (k - 1));
        T r = rhs[id];
        if (ft) r = m.g.dzC(k) * r;
        a[id] = cplx((double)r, 0.0);
    }
    if (ft) fourier_tridiagonal_poisson(m, a); else fft_poisson(m, a);
    for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
        size_t id = (size_t)(i - 1) + (size_t)Nx * ((size_t)(j - 1) + (size_t)Ny * (k - 1));
        phi(i, j, k) = (T)a[id].real();
    }
}

// ----------------------------------------------------------------------------------------
// ConjugateGradientPoissonSolver (conjugate_gradient_poisson_solver.jl:38-397) driving the
// ConjugateGradientSolver (conjugate_gradient_solver.jl:177-284).  Solves V∇²x = b with b = V·rhs.
// ----------------------------------------------------------------------------------------
template <class T> static bool cg_active(const Model<T> &m, int i, int j, int k) { return !IBP<T>(m).inactive_cell(i, j, k, false); }

// V∇²ᶜᶜᶜ (conjugate_gradient_poisson_solver.jl:62-66): δx(Ax ∂x c) + δy(Ay ∂y c) + δz(Az ∂z c)
// On an immersed grid both the inner ∂ and the outer δ are conditional (conditional_differences.jl):
// the gradient vanishes across a face next to an immersed cell, and so does the difference of the
// face fluxes around a cell bordered by an immersed inactive face.
template <class T> static T cg_laplacian(const Model<T> &m, const Fld<T> &c, int i, int j, int k) {
    const T idx = T(1) / m.g.dx, idy = T(1) / m.g.dy;
    IBP<T> P(m);
    auto zc = [&](int a, int b, int cc, int a2, int b2, int c2) {   // cell pair at (c, c, c)
        return m.ib && (P.imm_inactive(a, b, cc, 0, 0, 0) || P.imm_inactive(a2, b2, c2, 0, 0, 0));
    };
    auto zfp = [&](int a, int b, int cc, int a2, int b2, int c2, int lx, int ly, int lz) {   // face pair
        return m.ib && (P.imm_inactive(a, b, cc, lx, ly, lz) || P.imm_inactive(a2, b2, c2, lx, ly, lz));
    };
    auto gx = [&](int a) { return zc(a, j, k, a - 1, j, k) ? T(0) : (c(a, j, k) - c(a - 1, j, k)) * idx; };
    auto gy = [&](int b) { return zc(i, b, k, i, b - 1, k) ? T(0) : (c(i, b, k) - c(i, b - 1, k)) * idy; };
    auto gz = [&](int cc) { return zc(i, j, cc, i, j, cc - 1) ? T(0) : (c(i, j, cc) - c(i, j, cc - 1)) * (T(1) / m.g.dzF(cc)); };
    T tx = T(0), ty = T(0), tz = T(0);
    if (m.g.topo[0] != F && !zfp(i, j, k, i + 1, j, k, 1, 0, 0)) tx = m.Ax_c(k) * gx(i + 1) - m.Ax_c(k) * gx(i);
    if (m.g.topo[1] != F && !zfp(i, j, k, i, j + 1, k, 0, 1, 0)) ty = m.Ay_c(k) * gy(j + 1) - m.Ay_c(k) * gy(j);
    if (m.g.topo[2] != F && !zfp(i, j, k, i, j, k + 1, 0, 0, 1)) tz = m.Az() * gz(k + 1) - m.Az() * gz(k);
    return (tx + ty) + tz;
}

template <class T> static void cg_apply_laplacian(Model<T> &m, Fld<T> &out, Fld<T> &c) {
    fill_aux_halos(m, c);
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) out(i, j, k) = cg_laplacian(m, c, i, j, k);
}

// interior reductions in FT, summed in a fixed order (Field dot / norm / mean, field.jl:695-712)
template <class T, class F> static T cg_sum(const Model<T> &m, F f) {
    T s = T(0);
    for (int k = 1; k <= m.Nz(); ++k) for (int j = 1; j <= m.Ny(); ++j) for (int i = 1; i <= m.Nx(); ++i) s += f(i, j, k);
    return s;
}

// VolumeInverseNorm (conjugate_gradient_poisson_solver.jl:13-34): r ← r V⁻¹, ‖r‖, r ← r V
template <class T> static T cg_norm(Model<T> &m, Fld<T> &r) {
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
    for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) r(i, j, k) = r(i, j, k) * (T(1) / m.V_c(k));
    T n = std::sqrt(cg_sum<T>(m, [&](int i, int j, int k) { return r(i, j, k) * r(i, j, k); }));
    for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) r(i, j, k) = r(i, j, k) * m.V_c(k);
    return n;
}

// DiagonallyDominantPreconditioner (conjugate_gradient_poisson_solver.jl:256-296)
template <class T> static void cg_precondition_dd(Model<T> &m, Fld<T> &z, Fld<T> &r) {
    fill_aux_halos(m, r);
    auto Axm = [&](int, int, int k) { return m.Ax_c(k) * (T(1) / m.g.dx); };
    auto Aym = [&](int, int, int k) { return m.Ay_c(k) * (T(1) / m.g.dy); };
    auto Azm = [&](int, int, int k) { return m.Az() * (T(1) / m.g.dzF(k)); };
    auto Azp = [&](int, int, int k) { return m.Az() * (T(1) / m.g.dzF(k + 1)); };
    // a Flat direction has no coupling (its δ operators vanish)
    const bool fx = m.g.topo[0] == F, fy = m.g.topo[1] == F, fz = m.g.topo[2] == F;
    auto Ac = [&](int i, int j, int k) {
        T ax = fx ? T(0) : Axm(i, j, k), ay = fy ? T(0) : Aym(i, j, k);
        T azm = fz ? T(0) : Azm(i, j, k), azp = fz ? T(0) : Azp(i, j, k);
        return ((((-ax - ax) - ay) - ay) - azm) - azp;
    };
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
        T acc = r(i, j, k);
        if (!fx) {
            acc = acc - Axm(i, j, k) / Ac(i - 1, j, k) * r(i - 1, j, k);
            acc = acc - Axm(i + 1, j, k) / Ac(i + 1, j, k) * r(i + 1, j, k);
        }
        if (!fy) {
            acc = acc - Aym(i, j, k) / Ac(i, j - 1, k) * r(i, j - 1, k);
            acc = acc - Aym(i, j + 1, k) / Ac(i, j + 1, k) * r(i, j + 1, k);
        }
        if (!fz) {
            acc = acc - Azm(i, j, k) / Ac(i, j, k - 1) * r(i, j, k - 1);
            acc = acc - Azp(i, j, k) / Ac(i, j, k + 1) * r(i, j, k + 1);
        }
        z(i, j, k) = (T(1) / std::abs(Ac(i, j, k)) * acc) * (T)(cg_active(m, i, j, k) ? 1 : 0);
    }
}

// FFT-based preconditioner (conjugate_gradient_poisson_solver.jl:169-211): solve with rhs r·V⁻¹;
// the Fourier–tridiagonal solve!(x, solver, b) multiplies its source by Δzᶜ
// (fourier_tridiagonal_poisson_solver.jl:231-290), so both are the exact inverse of V∇².
template <class T> static void cg_precondition_fft(Model<T> &m, Fld<T> &z, Fld<T> &r) {
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
    std::vector<cplx> a((size_t)Nx * Ny * Nz);
    for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
        T v = r(i, j, k) * (T(1) / m.V_c(k));
        if (m.g.stretched) v = v * m.g.dzC(k);
        a[(size_t)(i - 1) + (size_t)Nx * ((size_t)(j - 1) + (size_t)Ny * (k - 1))] = cplx((double)v, 0.0);
    }
    if (m.g.stretched) fourier_tridiagonal_poisson(m, a); else fft_poisson(m, a);
    for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i)
        z(i, j, k) = (T)a[(size_t)(i - 1) + (size_t)Nx * ((size_t)(j - 1) + (size_t)Ny * (k - 1))].real();
}

// ColumnwiseTridiagonalPreconditioner (conjugate_gradient_poisson_solver.jl:298-397): per column the
// vertical part of V∇² with its diagonal scaled by 1 + 1/100 (identity rows for inactive and
// vertically isolated cells), solved by the batched Thomas sweep of batched_tridiagonal_solver.jl
// (solve_batched_tridiagonal_system_z!: forward sweep with the 10·eps pivot guard that keeps the
// output's previous value, then back substitution).
template <class T> static void cg_precondition_ct(Model<T> &m, Fld<T> &z, Fld<T> &r) {
    IBP<T> P(m);
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
    auto inact = [&](int i, int j, int k) { return P.inactive_cell(i, j, k, false); };
    auto azf = [&](int k) { return m.Az() * (T(1) / m.g.dzF(k)); };
    auto off = [&](int i, int j, int k) {
        const T ap = inact(i, j, k + 1) ? T(0) : azf(k + 1);
        return inact(i, j, k) ? T(0) : ap;
    };
    auto diag = [&](int i, int j, int k) {
        const bool s = inact(i, j, k), b = inact(i, j, k - 1), a = inact(i, j, k + 1);
        const T am = b ? T(0) : azf(k), ap = a ? T(0) : azf(k + 1);
        const T eps = T(1) / T(100);
        return (s || (b && a)) ? T(1) : -(am + ap) * (T(1) + eps);
    };
    const T guard = T(10) * std::numeric_limits<T>::epsilon();
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
        std::vector<T> t(Nz + 1);
        T beta = diag(i, j, 1);
        z(i, j, 1) = r(i, j, 1) / beta;
        for (int k = 2; k <= Nz; ++k) {
            const T c = off(i, j, k - 1), bk = diag(i, j, k), a = off(i, j, k - 1);
            t[k] = c / beta;
            beta = bk - a * t[k];
            const T phis = (r(i, j, k) - a * z(i, j, k - 1)) / beta;
            if (std::abs(beta) > guard) z(i, j, k) = phis;
        }
        for (int k = Nz - 1; k >= 1; --k) z(i, j, k) -= t[k + 1] * z(i, j, k + 1);
    }
}

// the reference's default: the FFT-based solver of the underlying grid on an immersed grid, the
// diagonally dominant preconditioner otherwise (conjugate_gradient_poisson_solver.jl:141-147)
template <class T> static int cg_preconditioner_kind(const Model<T> &m) {
    int pk = m.cfg.cg_preconditioner;
    if (pk == OM_PRECOND_DEFAULT) pk = m.ib ? OM_PRECOND_FFT : OM_PRECOND_DIAGONALLY_DOMINANT;
    return pk;
}

// solve!(x, ::ConjugateGradientSolver, b) (conjugate_gradient_solver.jl:177-284) with the zero-mean
// gauge (conjugate_gradient_poisson_solver.jl:87-96); x holds the initial guess.
template <class T> static void cg_solve(Model<T> &m, Fld<T> &x, Fld<T> &b) {
    Fld<T> &r = m.cg_r, &p = m.cg_p, &q = m.cg_q, &z = m.cg_z;
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
    const T seps = std::sqrt(std::numeric_limits<T>::epsilon());
    const T reltol = m.cfg.cg_reltol < 0 ? seps : (T)m.cfg.cg_reltol;
    const T abstol = m.cfg.cg_abstol < 0 ? seps : (T)m.cfg.cg_abstol;
    const int64_t maxiter = m.cfg.cg_maxiter > 0 ? m.cfg.cg_maxiter : (int64_t)Nx * Ny * Nz;
    const int pk = cg_preconditioner_kind(m);
    auto interior = [&](auto fn) {
        for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) fn(i, j, k);
    };
    int64_t iteration = 0;
    cg_apply_laplacian(m, q, x);
    interior([&](int i, int j, int k) { r(i, j, k) = b(i, j, k) - q(i, j, k); });
    const T res0 = cg_norm(m, r);
    const T tol = std::max(reltol * res0, abstol);
    T rho_prev = T(0);
    T res = cg_norm(m, r);
    while (iteration < maxiter && !(res <= tol)) {
        if (pk == OM_PRECOND_DIAGONALLY_DOMINANT) cg_precondition_dd(m, z, r);
        else if (pk == OM_PRECOND_FFT) cg_precondition_fft(m, z, r);
        else if (pk == OM_PRECOND_COLUMNWISE_TRIDIAGONAL) cg_precondition_ct(m, z, r);
        else interior([&](int i, int j, int k) { z(i, j, k) = r(i, j, k); });
        const T rho = cg_sum<T>(m, [&](int i, int j, int k) { return z(i, j, k) * r(i, j, k); });
        if (iteration == 0) interior([&](int i, int j, int k) { p(i, j, k) = z(i, j, k); });
        else {
            const T beta = rho / rho_prev;
            interior([&](int i, int j, int k) { p(i, j, k) = z(i, j, k) + beta * p(i, j, k); });
        }
        cg_apply_laplacian(m, q, p);
        const T alpha = rho / cg_sum<T>(m, [&](int i, int j, int k) { return p(i, j, k) * q(i, j, k); });
        interior([&](int i, int j, int k) {
            x(i, j, k) += alpha * p(i, j, k);
            r(i, j, k) -= alpha * q(i, j, k);
        });
        const T n = (T)(Nx * (double)Ny * Nz);
        const T mx = cg_sum<T>(m, [&](int i, int j, int k) { return x(i, j, k); }) / n;
        const T mr = cg_sum<T>(m, [&](int i, int j, int k) { return r(i, j, k); }) / n;
        interior([&](int i, int j, int k) {
            const T act = (T)(cg_active(m, i, j, k) ? 1 : 0);
            x(i, j, k) = (x(i, j, k) - mx) * act;
            r(i, j, k) = (r(i, j, k) - mr) * act;
        });
        iteration += 1;
        rho_prev = rho;
        res = cg_norm(m, r);
    }
    m.cg_iterations = (int)iteration;
    m.cg_residual = (double)res;
    // ‖V⁻¹(e − ē)‖ with e = b − A x recomputed in double precision (the zero-mean gauge of the iteration)
    fill_aux_halos(m, x);
    std::vector<double> ev((size_t)Nx * Ny * Nz);
    double esum = 0;
    interior([&](int i, int j, int k) {
        const double idx = 1.0 / (double)m.g.dx, idy = 1.0 / (double)m.g.dy;
        auto X = [&](int a, int b2, int c) { return (double)x(a, b2, c); };
        // the same conditional differences as cg_laplacian on an immersed grid
        IBP<T> P(m);
        auto zc = [&](int a, int b2, int c) { return m.ib && (P.imm_inactive(i, j, k, 0, 0, 0) || P.imm_inactive(a, b2, c, 0, 0, 0)); };
        auto zfp = [&](int a, int b2, int c, int lx, int ly, int lz) {
            return m.ib && (P.imm_inactive(i, j, k, lx, ly, lz) || P.imm_inactive(a, b2, c, lx, ly, lz));
        };
        auto G = [&](bool z, double v) { return z ? 0.0 : v; };
        double lap = 0;
        if (m.g.topo[0] != F && !zfp(i + 1, j, k, 1, 0, 0))
            lap += (double)m.Ax_c(k) * (G(zc(i + 1, j, k), (X(i + 1, j, k) - X(i, j, k)) * idx) - G(zc(i - 1, j, k), (X(i, j, k) - X(i - 1, j, k)) * idx));
        if (m.g.topo[1] != F && !zfp(i, j + 1, k, 0, 1, 0))
            lap += (double)m.Ay_c(k) * (G(zc(i, j + 1, k), (X(i, j + 1, k) - X(i, j, k)) * idy) - G(zc(i, j - 1, k), (X(i, j, k) - X(i, j - 1, k)) * idy));
        if (m.g.topo[2] != F && !zfp(i, j, k + 1, 0, 0, 1))
            lap += (double)m.Az() * (G(zc(i, j, k + 1), (X(i, j, k + 1) - X(i, j, k)) / (double)m.g.dzF(k + 1)) -
                                     G(zc(i, j, k - 1), (X(i, j, k) - X(i, j, k - 1)) / (double)m.g.dzF(k)));
        const double e = (double)b(i, j, k) - lap;
        ev[(size_t)(i - 1) + (size_t)Nx * ((size_t)(j - 1) + (size_t)Ny * (k - 1))] = e;
        esum += e;
    });
    const double emean = esum / ((double)Nx * Ny * Nz);
    double s = 0;
    interior([&](int i, int j, int k) {
        const double e = (ev[(size_t)(i - 1) + (size_t)Nx * ((size_t)(j - 1) + (size_t)Ny * (k - 1))] - emean) / (double)m.V_c(k);
        s += e * e;
    });
    m.cg_true_residual = std::sqrt(s);
}

// pressure_correction.jl:6-20
template <class T> static void compute_pressure_correction(Model<T> &m, T dt) {
    mask_immersed(m, m.U[0], 1, 0, 0);
    mask_immersed(m, m.U[1], 0, 1, 0);
    mask_immersed(m, m.U[2], 0, 0, 1);
    for (int f = 0; f < 3; ++f) fill_prognostic_halos(m, f, true);
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
    if (m.cfg.pressure_solver == OM_SOLVER_CONJUGATE_GRADIENT) {
        // solve_for_pressure.jl:149-161: warm start p·Δt★, rhs = active·δ·V
        const T eps = std::numeric_limits<T>::epsilon();
        const T dts = std::isfinite(dt) ? std::max(eps, dt) : T(0);
        for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) m.pNHS(i, j, k) *= dts;
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i)
            m.cg_b(i, j, k) = (T)(cg_active(m, i, j, k) ? 1 : 0) * divergence(m, i, j, k) * m.V_c(k);
        cg_solve(m, m.pNHS, m.cg_b);
        fill_aux_halos(m, m.pNHS);
        return;
    }
    std::vector<T> rhs((size_t)Nx * Ny * Nz);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i)
        rhs[(size_t)(i - 1) + (size_t)Nx * ((size_t)(j - 1) + (size_t)Ny * (k - 1))] = divergence(m, i, j, k);
    solve_pressure_field(m, rhs, m.pNHS);
    fill_aux_halos(m, m.pNHS);
}

// pressure_correction.jl:67-106
template <class T> static void make_pressure_correction(Model<T> &m, T dt) {
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
    Fld<T> &p = m.pNHS;
    IBP<T> P(m);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
        auto zd = [&](int i2, int j2, int k2) { return m.ib && (P.imm_inactive(i, j, k, 0, 0, 0) || P.imm_inactive(i2, j2, k2, 0, 0, 0)); };
        if (m.g.topo[0] != F) m.U[0](i, j, k) -= (zd(i - 1, j, k) ? T(0) : p(i, j, k) - p(i - 1, j, k)) * (T(1) / m.g.dx);
        if (m.g.topo[1] != F) m.U[1](i, j, k) -= (zd(i, j - 1, k) ? T(0) : p(i, j, k) - p(i, j - 1, k)) * (T(1) / m.g.dy);
        if (m.g.topo[2] != F) m.U[2](i, j, k) -= (zd(i, j, k - 1) ? T(0) : p(i, j, k) - p(i, j, k - 1)) * (T(1) / m.g.dzF(k));
    }
    T eps = std::numeric_limits<T>::epsilon();
    T dtp = std::max(eps, dt);
    for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) p(i, j, k) /= dtp;
}

// mask_immersed_field! (mask_immersed_field.jl:65-80): zero on the immersed periphery of the field's
// location, over the field's interior indices
template <class T> static void mask_immersed(Model<T> &m, Fld<T> &c, int lx, int ly, int lz) {
    if (!m.ib) return;
    IBP<T> P(m);
    const int khi = (lz && m.bz()) ? m.Nz() + 1 : m.Nz();
    for (int k = 1; k <= khi; ++k) for (int j = 1; j <= m.Ny(); ++j) for (int i = 1; i <= m.Nx(); ++i)
        if (P.imm_peripheral(i, j, k, lx, ly, lz)) c(i, j, k) = T(0);
}

// update_nonhydrostatic_model_state.jl:22-62
template <class T> static void update_state(Model<T> &m) {
    for (int n = 0; n < m.nt; ++n) mask_immersed(m, m.U[3 + n], 0, 0, 0);
    fill_all_prognostic_halos(m, false);
    compute_closure_fields(m);
    update_hydrostatic_pressure(m);
    compute_tendencies(m);
}

// runge_kutta_3.jl:196-204 + nonhydrostatic_rk3_substep.jl:31-63
template <class T> static void rk3_substep(Model<T> &m, double Dt, double gamma, double zeta, bool first) {
    T dt = (T)Dt;
    T g = (T)gamma, z = (T)zeta;
    T dtau = first ? (T)(Dt * (double)g) : (T)(Dt * ((double)g + (double)z));
    compute_flux_bc_tendencies(m);
    const int Nx = m.Nx(), Ny = m.Ny(), Nz = m.Nz();
    for (int f = 0; f < 3 + m.nt; ++f) {
        int klo = 1;
        if (f == 2 && m.bz() && Nz > 1) klo = 2;   // exclude_periphery: w skips the bottom boundary face
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = klo; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
            if (first) m.U[f](i, j, k) += dt * g * m.Gn[f](i, j, k);
            else m.U[f](i, j, k) += dt * (g * m.Gn[f](i, j, k) + z * m.Gm[f](i, j, k));
        }
    }
    compute_pressure_correction(m, dtau);
    make_pressure_correction(m, dtau);
}

template <class T> static void cache_previous_tendencies(Model<T> &m) {
    for (int f = 0; f < 3 + m.nt; ++f) m.Gm[f].d = m.Gn[f].d;
}

template <class T> static void time_step(Model<T> &m, double Dt) {
    const T g1 = (T)(8.0 / 15), g2 = (T)(5.0 / 12), g3 = (T)(3.0 / 4), z2 = (T)(-17.0 / 60), z3 = (T)(-5.0 / 12);
    if (m.iteration == 0) {
        m.last_dt = Dt;
        m.last_stage_dt = Dt * (double)g1;
        update_state(m);
    }
    double tn1 = m.time + Dt;
    rk3_substep(m, Dt, g1, 0, true);
    cache_previous_tendencies(m);
    m.time += Dt * (double)g1; m.stage = 2; m.last_stage_dt = Dt * (double)g1;
    update_state(m);
    rk3_substep(m, Dt, g2, z2, false);
    cache_previous_tendencies(m);
    m.time += Dt * ((double)g2 + (double)z2); m.stage = 3; m.last_stage_dt = Dt * ((double)g2 + (double)z2);
    update_state(m);
    rk3_substep(m, Dt, g3, z3, false);
    cache_previous_tendencies(m);
    double corrected = tn1 - m.time;
    m.time += corrected; m.iteration += 1; m.stage = 1; m.last_dt = Dt; m.last_stage_dt = corrected;
    update_state(m);
}

template <class T> static Fld<T> *field_by_id(Model<T> &m, int id) {
    if (id >= 0 && id < 3 + m.nt) return &m.U[id];
    if (id >= OM_FIELD_GN && id < OM_FIELD_GN + 3 + m.nt) return &m.Gn[id - OM_FIELD_GN];
    if (id >= OM_FIELD_GM && id < OM_FIELD_GM + 3 + m.nt) return &m.Gm[id - OM_FIELD_GM];
    if (id == OM_FIELD_PNHS) return &m.pNHS;
    if (id == OM_FIELD_PHY) return &m.pHY;
    if (id == OM_FIELD_NU) return &m.nu;
    if (id >= OM_FIELD_KAPPA0 && id < OM_FIELD_KAPPA0 + m.nt) return &m.kap[id - OM_FIELD_KAPPA0];
    return nullptr;
}

// interior extent of a field: w on Bounded z has Nz+1 faces
template <class T> static void interior_extent(const Model<T> &m, int id, int lo[3], int hi[3]) {
    for (int d = 0; d < 3; ++d) { lo[d] = 1; hi[d] = m.g.N[d]; }
    int base = id % 100;
    if (id < 300 && base == 2 && m.bz()) hi[2] = m.g.N[2] + 1;
}

}  // namespace orc

// ========================================================================================
// C interface for the ctypes harness (tests/, bench.py's cpu_baseline leg)
// ========================================================================================
using namespace orc;

struct orc_model {
    int precision;
    void *impl;
};

template <class T> static Model<T> *mk(const om_config_t *c) {
    auto *m = new Model<T>();
    m->cfg = *c;
    if (c->z_faces) { m->zfaces_copy.assign(c->z_faces, c->z_faces + c->N[2] + 1); m->cfg.z_faces = m->zfaces_copy.data(); }
    build_grid(m->g, &m->cfg);
    m->nt = c->n_tracers;
    int N[3] = {m->g.N[0], m->g.N[1], m->g.N[2]}, H[3] = {m->g.H[0], m->g.H[1], m->g.H[2]};
    for (int f = 0; f < 3 + m->nt; ++f) { m->U[f].init(N, H); m->Gn[f].init(N, H); m->Gm[f].init(N, H); }
    m->pNHS.init(N, H); m->pHY.init(N, H); m->nu.init(N, H); m->rhs.init(N, H);
    if (c->pressure_solver == OM_SOLVER_CONJUGATE_GRADIENT) {
        m->cg_r.init(N, H); m->cg_p.init(N, H); m->cg_q.init(N, H); m->cg_z.init(N, H); m->cg_b.init(N, H);
    }
    for (int n = 0; n < m->nt; ++n) m->kap[n].init(N, H);
    if (c->bottom_height) {
        // numerical bottom: the face above the last cell whose centre lies at or below z_b
        // (grid_fitted_bottom.jl:118-139, CenterImmersedCondition)
        m->ib = true;
        m->kbot_.assign((size_t)N[0] * N[1], 0);
        for (int j = 1; j <= N[1]; ++j) for (int i = 1; i <= N[0]; ++i) {
            const double zb = c->bottom_height[(size_t)(i - 1) + (size_t)N[0] * (j - 1)];
            int kb = 0;
            using FT = decltype(m->g.dx);
            for (int k = 1; k <= N[2]; ++k) if ((FT)m->g.zC(k) <= (FT)zb) kb = k;
            m->kbot_[(size_t)(i - 1) + (size_t)N[0] * (j - 1)] = kb;
        }
    }
    int rule = c->hydrostatic_pressure;
    if (rule < 0) m->has_pHY = (c->topology[2] != P) && (c->topology[2] != F) && c->buoyancy != OM_BUOYANCY_NONE;
    else m->has_pHY = rule == 1;
    return m;
}

#define DISPATCH(h, ...)                                                                    \
    do {                                                                                    \
        if ((h)->precision == 8) { auto *m = (Model<double> *)(h)->impl; __VA_ARGS__; }     \
        else { auto *m = (Model<float> *)(h)->impl; __VA_ARGS__; }                          \
    } while (0)

extern "C" {

void *orc_create(const om_config_t *c) {
    if (!c) return nullptr;
    for (int d = 0; d < 2; ++d) if (c->topology[d] == OM_BOUNDED) return nullptr;   // x, y: Periodic or Flat
    auto *h = new orc_model();
    h->precision = c->precision == 4 ? 4 : 8;
    if (h->precision == 8) h->impl = mk<double>(c); else h->impl = mk<float>(c);
    return h;
}

void orc_destroy(void *hp) {
    auto *h = (orc_model *)hp;
    if (!h) return;
    if (h->precision == 8) delete (Model<double> *)h->impl; else delete (Model<float> *)h->impl;
    delete h;
}

// copy the interior (or the whole padded parent when with_halos) of a field to/from host (FP64)
int orc_set_field(void *hp, int id, const double *src) {
    auto *h = (orc_model *)hp;
    int rc = 0;
    DISPATCH(h, {
        auto *f = field_by_id(*m, id);
        if (!f) { rc = -1; break; }
        int lo[3], hi[3]; interior_extent(*m, id, lo, hi);
        size_t t = 0;
        for (int k = lo[2]; k <= hi[2]; ++k) for (int j = lo[1]; j <= hi[1]; ++j) for (int i = lo[0]; i <= hi[0]; ++i)
            (*f)(i, j, k) = (std::decay_t<decltype(f->d[0])>)src[t++];
    });
    return rc;
}

int orc_get_field(void *hp, int id, double *dst) {
    auto *h = (orc_model *)hp;
    int rc = 0;
    DISPATCH(h, {
        auto *f = field_by_id(*m, id);
        if (!f) { rc = -1; break; }
        int lo[3], hi[3]; interior_extent(*m, id, lo, hi);
        size_t t = 0;
        for (int k = lo[2]; k <= hi[2]; ++k) for (int j = lo[1]; j <= hi[1]; ++j) for (int i = lo[0]; i <= hi[0]; ++i)
            dst[t++] = (double)(*f)(i, j, k);
    });
    return rc;
}

// parent-array access with halos: ranges lo..hi inclusive in reference indices
int orc_get_field_range(void *hp, int id, const int lo[3], const int hi[3], double *dst) {
    auto *h = (orc_model *)hp;
    int rc = 0;
    DISPATCH(h, {
        auto *f = field_by_id(*m, id);
        if (!f) { rc = -1; break; }
        size_t t = 0;
        for (int k = lo[2]; k <= hi[2]; ++k) for (int j = lo[1]; j <= hi[1]; ++j) for (int i = lo[0]; i <= hi[0]; ++i)
            dst[t++] = (double)(*f)(i, j, k);
    });
    return rc;
}
int orc_set_field_range(void *hp, int id, const int lo[3], const int hi[3], const double *src) {
    auto *h = (orc_model *)hp;
    int rc = 0;
    DISPATCH(h, {
        auto *f = field_by_id(*m, id);
        if (!f) { rc = -1; break; }
        size_t t = 0;
        for (int k = lo[2]; k <= hi[2]; ++k) for (int j = lo[1]; j <= hi[1]; ++j) for (int i = lo[0]; i <= hi[0]; ++i)
            (*f)(i, j, k) = (std::decay_t<decltype(f->d[0])>)src[t++];
    });
    return rc;
}

void orc_fill_halos(void *hp, int fill_normal) { auto *h = (orc_model *)hp; DISPATCH(h, fill_all_prognostic_halos(*m, fill_normal != 0)); }
void orc_update_state(void *hp) { auto *h = (orc_model *)hp; DISPATCH(h, update_state(*m)); }
void orc_compute_closure(void *hp) { auto *h = (orc_model *)hp; DISPATCH(h, compute_closure_fields(*m)); }
void orc_compute_hydrostatic(void *hp) { auto *h = (orc_model *)hp; DISPATCH(h, update_hydrostatic_pressure(*m)); }
void orc_compute_tendencies(void *hp) { auto *h = (orc_model *)hp; DISPATCH(h, compute_tendencies(*m)); }
void orc_rk3_substep(void *hp, double dt, double g, double z, int first) { auto *h = (orc_model *)hp; DISPATCH(h, rk3_substep(*m, dt, g, z, first != 0)); }
void orc_cache_previous_tendencies(void *hp) { auto *h = (orc_model *)hp; DISPATCH(h, cache_previous_tendencies(*m)); }
void orc_pressure_correction(void *hp, double dt) {
    auto *h = (orc_model *)hp;
    DISPATCH(h, { compute_pressure_correction(*m, (decltype(m->g.dx))dt); make_pressure_correction(*m, (decltype(m->g.dx))dt); });
}
void orc_time_step(void *hp, double dt, int n) { auto *h = (orc_model *)hp; for (int s = 0; s < n; ++s) DISPATCH(h, time_step(*m, dt)); }

// set!(model; ...) epilogue: halos of the set fields, update_state!, projection with Δt = 1, update_state!
void orc_set_finalize(void *hp, int enforce_incompressibility) {
    auto *h = (orc_model *)hp;
    DISPATCH(h, {
        fill_all_prognostic_halos(*m, true);
        update_state(*m);
        if (enforce_incompressibility) {
            compute_pressure_correction(*m, (decltype(m->g.dx))1);
            make_pressure_correction(*m, (decltype(m->g.dx))1);
            update_state(*m);
        }
    });
}

void orc_last_solve_info(void *hp, int *iterations, double *residual, double *true_residual) {
    auto *h = (orc_model *)hp;
    DISPATCH(h, { *iterations = m->cg_iterations; *residual = m->cg_residual; *true_residual = m->cg_true_residual; });
}

// cell_advection_timescale(model) (cell_advection_timescale.jl:14-35): minimum over interior cells
// of 1 / (|u|Δx⁻¹ + |v|Δy⁻¹ + |w|Δz⁻¹ᶜᶜᶠ) in FT, a Flat direction contributing 0
double orc_cell_advection_timescale(void *hp) {
    auto *h = (orc_model *)hp;
    double out = 0;
    DISPATCH(h, {
        using FT = decltype(m->g.dx);
        FT tau = std::numeric_limits<FT>::infinity();
        bool nan = false;
        for (int k = 1; k <= m->Nz(); ++k) for (int j = 1; j <= m->Ny(); ++j) for (int i = 1; i <= m->Nx(); ++i) {
            const FT ix = m->g.topo[0] == F ? FT(0) : std::abs(m->U[0](i, j, k)) * (FT(1) / m->g.dx);
            const FT iy = m->g.topo[1] == F ? FT(0) : std::abs(m->U[1](i, j, k)) * (FT(1) / m->g.dy);
            const FT iz = m->g.topo[2] == F ? FT(0) : std::abs(m->U[2](i, j, k)) * (FT(1) / m->g.dzF(k));
            const FT t = FT(1) / ((ix + iy) + iz);
            if (t != t) nan = true;
            else tau = std::min(tau, t);
        }
        out = nan ? std::numeric_limits<double>::quiet_NaN() : (double)tau;
    });
    return out;
}

// one application of the CG preconditioner `kind` (OM_PRECOND_*) to an interior residual r
void orc_cg_precondition(void *hp, int kind, const double *r, double *z) {
    auto *h = (orc_model *)hp;
    DISPATCH(h, {
        using FT = decltype(m->g.dx);
        int N[3] = {m->Nx(), m->Ny(), m->Nz()}, H[3] = {m->g.H[0], m->g.H[1], m->g.H[2]};
        Fld<FT> rr, zz;
        rr.init(N, H);
        zz.init(N, H);
        size_t t = 0;
        for (int k = 1; k <= N[2]; ++k) for (int j = 1; j <= N[1]; ++j) for (int i = 1; i <= N[0]; ++i) rr(i, j, k) = (FT)r[t++];
        if (kind == OM_PRECOND_DIAGONALLY_DOMINANT) cg_precondition_dd(*m, zz, rr);
        else if (kind == OM_PRECOND_FFT) cg_precondition_fft(*m, zz, rr);
        else if (kind == OM_PRECOND_COLUMNWISE_TRIDIAGONAL) cg_precondition_ct(*m, zz, rr);
        else zz = rr;
        t = 0;
        for (int k = 1; k <= N[2]; ++k) for (int j = 1; j <= N[1]; ++j) for (int i = 1; i <= N[0]; ++i) z[t++] = (double)zz(i, j, k);
    });
}

void orc_get_clock(void *hp, double *time, int64_t *it, int *stage) {
    auto *h = (orc_model *)hp;
    DISPATCH(h, { *time = m->time; *it = m->iteration; *stage = m->stage; });
}
// OpenMP threads of the oracle's later parallel regions started from the calling thread
void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
// restore_prognostic_state! of the clock (nonhydrostatic_model.jl:354-375, clock.jl:147-173)
void orc_set_clock(void *hp, double time, int64_t it, int stage, double last_dt, double last_stage_dt) {
    auto *h = (orc_model *)hp;
    DISPATCH(h, { m->time = time; m->iteration = it; m->stage = stage; m->last_dt = last_dt; m->last_stage_dt = last_stage_dt; });
}

// ∇·U at cell centres (interior)
void orc_divergence(void *hp, double *out) {
    auto *h = (orc_model *)hp;
    DISPATCH(h, {
        size_t t = 0;
        for (int k = 1; k <= m->Nz(); ++k) for (int j = 1; j <= m->Ny(); ++j) for (int i = 1; i <= m->Nx(); ++i)
            out[t++] = (double)divergence(*m, i, j, k);
    });
}

// solve the model's pressure Poisson problem for an arbitrary interior rhs (FP64 in/out)
void orc_poisson_solve(void *hp, const double *rhs, double *phi) {
    auto *h = (orc_model *)hp;
    DISPATCH(h, {
        using TT = decltype(m->g.dx);
        size_t n = (size_t)m->Nx() * m->Ny() * m->Nz();
        std::vector<TT> r(n);
        for (size_t t = 0; t < n; ++t) r[t] = (TT)rhs[t];
        solve_pressure_field(*m, r, m->rhs);
        size_t t = 0;
        for (int k = 1; k <= m->Nz(); ++k) for (int j = 1; j <= m->Ny(); ++j) for (int i = 1; i <= m->Nx(); ++i)
            phi[t++] = (double)m->rhs(i, j, k);
    });
}

// grid metrics for tests: dzC and dzF for k = 1-Hz .. Nz+Hz (2Hz+Nz values each)
void orc_grid_z(void *hp, double *dzc, double *dzf, double *zc, double *zf) {
    auto *h = (orc_model *)hp;
    DISPATCH(h, {
        int Hz = m->g.H[2], Nz = m->g.N[2], t = 0;
        for (int k = 1 - Hz; k <= Nz + Hz; ++k) {
            dzc[t] = (double)m->g.dzC(k); dzf[t] = (double)m->g.dzF(k); zc[t] = m->g.zC(k); zf[t] = m->g.zF(k); ++t;
        }
    });
}

// 1D reconstructions for tests (weno5 on 6 values, weno3 on 4 values, centered4 on 4 values)
double orc_weno5(const double *S, int left, int fp32) {
    if (fp32) { Coeffs<float> K; float s[6]; for (int i = 0; i < 6; ++i) s[i] = (float)S[i]; return weno5<float>(K, s, left != 0); }
    Coeffs<double> K; return weno5<double>(K, S, left != 0);
}
double orc_weno3(const double *S, int left, int fp32) {
    if (fp32) { Coeffs<float> K; float s[4]; for (int i = 0; i < 4; ++i) s[i] = (float)S[i]; return weno3<float>(K, s, left != 0); }
    Coeffs<double> K; return weno3<double>(K, S, left != 0);
}
// Centered(4) symmetric interpolation at face i from S = (ψ[i-2], ψ[i-1], ψ[i], ψ[i+1])
double orc_centered4(const double *S, int fp32) {
    if (fp32) { Coeffs<float> K; float s[4]; for (int i = 0; i < 4; ++i) s[i] = (float)S[i];
                return rec_symmetric(K, Z4, [&](int o) { return s[o + 2]; }); }
    Coeffs<double> K;
    return rec_symmetric(K, Z4, [&](int o) { return S[o + 2]; });
}
void orc_weno5_smoothness(const double *S, int left, int fp32, double *beta_out, double *weights_out) {
    auto run = [&](auto tag) {
        using TT = decltype(tag);
        Coeffs<TT> K;
        TT s[6]; for (int i = 0; i < 6; ++i) s[i] = (TT)S[i];
        TT ps[3][3];
        if (left) { for (int r = 0; r < 3; ++r) for (int q = 0; q < 3; ++q) ps[r][q] = s[2 - r + q]; }
        else { for (int r = 0; r < 3; ++r) for (int q = 0; q < 3; ++q) ps[r][q] = s[3 + r - q]; }
        TT beta[3], alpha[3];
        for (int r = 0; r < 3; ++r) beta[r] = beta5<TT>(ps[r], r);
        TT tau = std::abs(beta[0] - beta[2]);
        for (int r = 0; r < 3; ++r) { TT q = tau / (beta[r] + K.eps); alpha[r] = K.Cs5[r] * (TT(1) + q * q); }
        TT sinv = TT(1) / ((alpha[0] + alpha[1]) + alpha[2]);
        for (int r = 0; r < 3; ++r) { beta_out[r] = (double)beta[r]; weights_out[r] = (double)(alpha[r] * sinv); }
    };
    if (fp32) run(float(0)); else run(double(0));
}
void orc_coefficients(int fp32, double *w5, double *w3, double *c4, double *c2) {
    auto run = [&](auto tag) {
        using TT = decltype(tag);
        Coeffs<TT> K;
        for (int r = 0; r < 3; ++r) for (int q = 0; q < 3; ++q) w5[3 * r + q] = (double)K.w5[r][q];
        for (int r = 0; r < 2; ++r) for (int q = 0; q < 2; ++q) w3[2 * r + q] = (double)K.w3[r][q];
        for (int q = 0; q < 4; ++q) c4[q] = (double)K.c4[q];
        for (int q = 0; q < 2; ++q) c2[q] = (double)K.c2[q];
    };
    if (fp32) run(float(0)); else run(double(0));
}

// batched Thomas solve of general tridiagonal systems (batched_tridiagonal_solver.jl:211-243)
void orc_thomas(int n, int nbatch, const double *a, const double *b, const double *c, const double *f, double *x) {
    for (int s = 0; s < nbatch; ++s) {
        const double *bb = b + (size_t)s * n, *ff = f + (size_t)s * n;
        double *xx = x + (size_t)s * n;
        std::vector<double> t(n);
        double beta = bb[0];
        xx[0] = ff[0] / beta;
        for (int k = 1; k < n; ++k) {
            t[k] = c[(size_t)s * (n - 1) + k - 1] / beta;
            beta = bb[k] - a[(size_t)s * (n - 1) + k - 1] * t[k];
            if (std::abs(beta) > 10 * std::numeric_limits<double>::epsilon())
                xx[k] = (ff[k] - a[(size_t)s * (n - 1) + k - 1] * xx[k - 1]) / beta;
            else xx[k] = 0;
        }
        for (int k = n - 2; k >= 0; --k) xx[k] -= t[k + 1] * xx[k + 1];
    }
}

}  // extern "C"
