// om_diagnostics.hip — device reductions behind the time-step wizard and the NaN checker:
//   cell_advection_timescale(model)  src/Models/NonhydrostaticModels/NonhydrostaticModels.jl:93-97,
//                                    src/Advection/cell_advection_timescale.jl:14-35
//   maximum / minimum / hasnan of a field's data (parent or interior), as used by
//   cell_diffusion_timescale (src/TurbulenceClosures/turbulence_closure_diagnostics.jl:50-70) and
//   NaNChecker (src/Diagnostics/nan_checker.jl:37-38)
//
// Each reduction is two launches of fixed shape (per-block partials, then one block over the
// partials), so results do not depend on scheduling.  max and min are exact operations, so the
// advection timescale min over cells of 1 / (|u|Δx⁻¹ + |v|Δy⁻¹ + |w|Δz⁻¹) equals 1 / (max over
// cells of the inverse timescale), the reciprocal being monotone; a NaN anywhere makes it NaN, as
// Julia's `minimum` does.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "om_kernels.h"

namespace om {

constexpr int kDiagBlocks = 1024;
constexpr int kDiagThreads = 256;

struct Stats {
    double mx, mn, nan;
};

__device__ __forceinline__ Stats stats_combine(Stats a, Stats b) {
    return Stats{fmax(a.mx, b.mx), fmin(a.mn, b.mn), a.nan + b.nan};
}

// block reduction of (max, min, NaN count); the result is valid in thread 0
__device__ __forceinline__ Stats block_stats(Stats v) {
    __shared__ double smx[kDiagThreads / 64], smn[kDiagThreads / 64], snan[kDiagThreads / 64];
    for (int o = 32; o > 0; o >>= 1) {
        Stats w{__shfl_down(v.mx, o, 64), __shfl_down(v.mn, o, 64), __shfl_down(v.nan, o, 64)};
        v = stats_combine(v, w);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) { smx[wid] = v.mx; smn[wid] = v.mn; snan[wid] = v.nan; }
    __syncthreads();
    Stats r{-INFINITY, INFINITY, 0.0};
    if (threadIdx.x == 0)
        for (int w = 0; w < kDiagThreads / 64; ++w) r = stats_combine(r, Stats{smx[w], smn[w], snan[w]});
    __syncthreads();
    return r;
}

// inverse advective timescale of interior cell (i, j, k): |u[i]|/Δxᶠᶜᶜ + |v[j]|/Δyᶜᶠᶜ + |w[k]|/Δzᶜᶜᶠ,
// a Flat direction contributing 0
template <class T> __global__ void __launch_bounds__(kDiagThreads) k_adv_inv(Ctx<T> C, double *partial) {
#pragma clang fp contract(off)
    const DevGrid<T> &g = C.g;
    const int64_t n = (int64_t)g.Nx * g.Ny * g.Nz;
    Stats acc{-INFINITY, INFINITY, 0.0};
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.Nx);
        const int64_t r = t / g.Nx;
        const int j = (int)(r % g.Ny), k = (int)(r / g.Ny);
        const int64_t o = idx3(i, j, k, g.sy, g.sz);
        const T ix = C.flx() ? T(0) : fabs(C.s.U[0][o]) * g.idx;
        const T iy = C.fly() ? T(0) : fabs(C.s.U[1][o]) * g.idy;
        const T iz = C.flz() ? T(0) : fabs(C.s.U[2][o]) * g.idzf[k];
        const T inv = (ix + iy) + iz;
        if (inv != inv) acc.nan += 1.0;
        else acc.mx = fmax(acc.mx, (double)inv);
    }
    const Stats s = block_stats(acc);
    if (threadIdx.x == 0) { partial[blockIdx.x] = s.mx; partial[kDiagBlocks + blockIdx.x] = s.nan; }
}

// statistics of a field over its parent array (halos included) or its interior
template <class T> __global__ void __launch_bounds__(kDiagThreads) k_field_stats(DevGrid<T> g, const T *origin, int with_halos,
                                                                               int face_z, double *partial) {
    Stats acc{-INFINITY, INFINITY, 0.0};
    const int nx = with_halos ? g.Nx + 2 * g.Hx : g.Nx;
    const int ny = with_halos ? g.Ny + 2 * g.Hy : g.Ny;
    const int nz = with_halos ? g.Nz + 2 * g.Hz + (face_z ? 1 : 0) : g.Nz + (face_z ? 1 : 0);
    const int i0 = with_halos ? -g.Hx : 0, j0 = with_halos ? -g.Hy : 0, k0 = with_halos ? -g.Hz : 0;
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx);
        const int64_t r = t / nx;
        const int j = (int)(r % ny), k = (int)(r / ny);
        const double v = (double)origin[idx3(i + i0, j + j0, k + k0, g.sy, g.sz)];
        if (v != v) acc.nan += 1.0;
        else { acc.mx = fmax(acc.mx, v); acc.mn = fmin(acc.mn, v); }
    }
    const Stats s = block_stats(acc);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s.mx;
        partial[kDiagBlocks + blockIdx.x] = s.mn;
        partial[2 * kDiagBlocks + blockIdx.x] = s.nan;
    }
}

// out[0..2] = (max, min, NaN count) over the partials, in a fixed order
static __global__ void __launch_bounds__(kDiagThreads) k_stats_finish(const double *partial, int have_min, double *out) {
    Stats acc{-INFINITY, INFINITY, 0.0};
    for (int b = threadIdx.x; b < kDiagBlocks; b += blockDim.x) {
        const Stats v{partial[b], have_min ? partial[kDiagBlocks + b] : INFINITY, partial[(have_min ? 2 : 1) * kDiagBlocks + b]};
        acc = stats_combine(acc, v);
    }
    const Stats s = block_stats(acc);
    if (threadIdx.x == 0) { out[0] = s.mx; out[1] = s.mn; out[2] = s.nan; }
}

template <class T> void launch_adv_timescale(const Ctx<T> &C, double *partial, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_adv_inv<T>, dim3(kDiagBlocks), dim3(kDiagThreads), 0, st, C, partial);
    hipLaunchKernelGGL(k_stats_finish, dim3(1), dim3(kDiagThreads), 0, st, (const double *)partial, 0, out);
}

template <class T> void launch_field_stats(const DevGrid<T> &g, const T *origin, int with_halos, int face_z, double *partial,
                                           double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_field_stats<T>, dim3(kDiagBlocks), dim3(kDiagThreads), 0, st, g, origin, with_halos, face_z, partial);
    hipLaunchKernelGGL(k_stats_finish, dim3(1), dim3(kDiagThreads), 0, st, (const double *)partial, 1, out);
}

#ifndef OM_ONLY_DOUBLE
// ----------------------------------------------------------------------------------------
// Allocation-free reductions.  `scratch` holds the per-block partials (3 × kDiagBlocks doubles)
// and the three results of one field-statistics reduction; it lives in the library's code object,
// so each GPU holds its own copy from load to process end and no call allocates.  In host code
// `scratch` names a host-side stand-in, never the device memory: the device address is taken once
// per device with hipGetSymbolAddress (with that device current) and only that address is handed
// to kernels and copies.  Every entry point that uses it on one GPU shares the one array, so calls
// that overlap on one GPU race.
// ----------------------------------------------------------------------------------------
constexpr int kScratchDoubles = 3 * kDiagBlocks + 3;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mutex;
static double *g_scratch_ptr[kMaxScratchDevices];

double *diag_scratch(int device, int *err) {
    *err = 0;
    if (device < 0 || device >= kMaxScratchDevices) { *err = -1; return nullptr; }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != device) e = hipSetDevice(device);
        void *p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) { *err = (int)e; return nullptr; }
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    return g_scratch_ptr[device];
}
int diag_scratch_doubles() { return kScratchDoubles; }
int diag_partial_doubles() { return 3 * kDiagBlocks; }
#endif

#define OM_DIAG_INSTANTIATE(T)                                                                              \
    template void launch_adv_timescale<T>(const Ctx<T> &, double *, double *, hipStream_t);                 \
    template void launch_field_stats<T>(const DevGrid<T> &, const T *, int, int, double *, double *, hipStream_t);
#ifndef OM_ONLY_DOUBLE
OM_DIAG_INSTANTIATE(float)
#endif
#ifndef OM_ONLY_FLOAT
OM_DIAG_INSTANTIATE(double)
#endif

}  // namespace om
