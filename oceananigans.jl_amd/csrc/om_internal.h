// om_internal.h — device-side views shared by the HIP kernels and the C-ABI host code.
//
// Memory layout (HBM): every 3D field is one dense block, x fastest, with the parent shape
// (Nx+2Hx) × (Ny+2Hy) × (Nz+2Hz+1).  All fields share the same strides, so one index serves
// Center and Face fields (a Center field simply never touches its last z plane).  The view
// pointer `p` addresses the first interior point; device indices are 0-based (i = 0 is the
// reference's i = 1), so halo points have negative indices or indices ≥ N.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/oceananigans_mi355x.h"

namespace om {

constexpr int kMaxTracers = OM_MAX_TRACERS;
constexpr int kMaxFields = OM_MAX_FIELDS;

template <class T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

// Grid description passed by value to every kernel.
template <class T> struct DevGrid {
    int Nx, Ny, Nz;
    int Hx, Hy, Hz;
    int tx, ty, tz;          // topology per axis
    int64_t sy, sz;          // strides (elements) of the y and z axes
    T dx, dy, idx, idy;      // uniform horizontal spacings and their reciprocals
    T Az, iAz;               // Δx·Δy
    const T *dzc, *dzf;      // Δzᶜ[k], Δzᶠ[k] for k = -Hz .. Nz+Hz (pointer offset so [k] is valid)
    const T *idzc, *idzf;    // reciprocals
    const double *idzf64;    // 1 / (double)Δzᶠ[k], IEEE-rounded (the FP64 operator sweeps' factor)
    int stretched;
    const int *kbot;         // immersed bottom: immersed cells per interior column (x fastest), or nullptr
    int kbot_top;            // the largest entry of kbot: no cell at or above this level is immersed
    // Levels per column (x fastest) on which an immersed-boundary rule can apply to a tendency or
    // closure evaluation (an immersed cell within kImmersedReach cells in x, y and z), 0 where none
    // can; elsewhere the evaluation equals that on the grid without the bottom.  nullptr without one.
    const int *kfix;
    int kfix_max;            // the largest entry of kfix
    // the (64 × 4-column block, level) pairs of the generic kernels' launch that hold a cell below
    // kfix: x = block x | block y << 16, y = level; the fix-up launches visit only these
    const int2 *fix_list;
    int fix_count;
    // per-level AMD ratios, 5 entries for each k = -1 .. Nz at (k + 1)·5: 2Δx/Δᶠz, Δᶠz/2Δx, 2Δy/Δᶠz,
    // Δᶠz/2Δy with Δᶠz = 2Δzᶜ(k), and δ² = 3 / (1/(2Δx)² + 1/(2Δy)² + 1/Δᶠz²), rounded in FT as the
    // per-cell expressions (anisotropic_minimum_dissipation.jl:161-191)
    const T *amdr;
};

template <class T> struct DevField {
    T *p;  // address of interior point (0,0,0)
};

// Physics parameters of the tendency / closure kernels.
template <class T> struct DevPhysics {
    int advection, closure, buoyancy, coriolis, hydrostatic;
    int n_tracers, b_tracer, T_tracer, S_tracer;
    T f, g, alpha, beta;
    T c0, c1, c2;            // closure parameters (AMD: Cν, Cκ; Smag: C, Cb, Pr; constant: ν, κ)
    int bc_bottom[kMaxFields], bc_top[kMaxFields];
    T bc_bottom_value[kMaxFields], bc_top_value[kMaxFields];
};

template <class T> struct StateView {
    T *U[kMaxFields];        // u, v, w, tracers (interior origin)
    T *Gn[kMaxFields];
    T *Gm[kMaxFields];
    T *nu;                   // νₑ
    T *kap[kMaxTracers];     // κₑ per tracer (AMD)
    T *pHY;                  // hydrostatic pressure anomaly
    // FP32 only (nullptr in FP64): pHY′(i) − pHY′(i−1) and pHY′(j) − pHY′(j−1) at each cell, formed as
    // column sums of the buoyancy's horizontal differences, which FP32 holds to its relative
    // precision where the difference of two rounded column sums would not
    T *dxpHY, *dypHY;
    T *pNHS;
    const T *btr, *Ttr, *Str;  // buoyancy tracer / temperature / salinity (resolved on the host)
};

__host__ __device__ inline int64_t idx3(int i, int j, int k, int64_t sy, int64_t sz) {
    return (int64_t)i + (int64_t)j * sy + (int64_t)k * sz;
}

}  // namespace om
