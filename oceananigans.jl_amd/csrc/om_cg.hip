// om_cg.hip — kernels of the preconditioned conjugate-gradient Poisson solver
// (ConjugateGradientPoissonSolver, src/Solvers/conjugate_gradient_poisson_solver.jl:38-397, driving
// ConjugateGradientSolver, src/Solvers/conjugate_gradient_solver.jl:177-284).
//
// The solver works on V∇²x = b.  The iterate x, the residual r, the right-hand side b, the search
// direction p and A·p are FP64 in both precisions; only the preconditioner (and its output z) runs
// in the model's precision.  In an FP32 model the iteration therefore converges below the FP32
// floor ε·κ(A) (a residual of ~1e-4 for an exact FP32 inverse on a 32³ grid), down to the
// requested tolerance; the pressure field receives x rounded to FP32.  Every inner product and norm
// accumulates in FP64 through a two-stage reduction with a fixed grid shape (per-block partial sums,
// then one block summing the partials in a fixed order), so a solve is bitwise reproducible.
#include <hip/hip_runtime.h>

#include <climits>
#include <limits>

#include "om_kernels.h"

namespace om {

constexpr int kRedBlocks = kCgRedBlocks;   // blocks and partial sums per value of a grid-wide reduction
constexpr int kRedThreads = 256;

// sum of v over the block, valid in thread 0 (64-wide waves, then across the 4 waves in LDS)
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double wsum[kRedThreads / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) wsum[wid] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kRedThreads / 64; ++w) s += wsum[w];
    __syncthreads();
    return s;
}

// The interior cells one thread visits.  A work unit is a 64-cell segment of one x row, taken by
// one wave with a cell per lane: the unit's row (j, k) comes from scalar arithmetic and its
// accesses are contiguous.  The waves of the grid take the units in turn, two per pass, so the
// second unit's loads need not wait for the first unit's stores.  f(i, j, k, o, t) receives the
// padded index o and the dense index t = i + Nx·(j + Ny·k).
template <class T, class F> __device__ __forceinline__ void for_cells(const DevGrid<T> &g, F &&f) {
    const int lane = threadIdx.x & 63;
    const int nseg = (g.Nx + 63) >> 6;
    const int units = nseg * g.Ny * g.Nz;
    const int nw = (int)gridDim.x * ((int)blockDim.x >> 6);
    const int w = __builtin_amdgcn_readfirstlane((int)blockIdx.x * ((int)blockDim.x >> 6) + ((int)threadIdx.x >> 6));
    auto visit = [&](int u) {
        const int row = u / nseg, seg = u - row * nseg;
        const int k = row / g.Ny, j = row - k * g.Ny;
        const int i = (seg << 6) + lane;
        if (i < g.Nx) f(i, j, k, idx3(i, j, k, g.sy, g.sz), (int64_t)i + (int64_t)g.Nx * row);
    };
    int u = w;
    for (; u + nw < units; u += 2 * nw) {
        visit(u);
        visit(u + nw);
    }
    if (u < units) visit(u);
}

template <class T> __device__ __forceinline__ T vol(const DevGrid<T> &g, int k) { return g.Az * g.dzc[k]; }

// 1 for an active cell, 0 for an immersed one
template <class T> __device__ __forceinline__ double active(const DevGrid<T> &g, int i, int j, int k) {
    return IB<T>(g).inactive_cell(i, j, k) ? 0.0 : 1.0;
}

// The immersed-boundary predicates of the 7-point operator at any level of one column, from the
// bottom indices of the column and of its four horizontal neighbours, loaded once (the rules of
// IB::iin, grid_fitted_immersed_boundaries.jl / immersed_boundary_grid.jl).  Without an immersed
// bottom every index is INT_MIN, which makes every immersed predicate false.
template <class T> struct ColumnMask {
    int c, xm, xp, ym, yp;
    int Nz;
    bool bounded;
    __device__ __forceinline__ ColumnMask(const DevGrid<T> &g, int i, int j) : Nz(g.Nz), bounded(g.tz == OM_BOUNDED) {
        if (g.kbot) {
            auto kb = [&](int ii, int jj) { return g.kbot[IB<T>::wrap(ii, g.Nx) + g.Nx * IB<T>::wrap(jj, g.Ny)]; };
            c = kb(i, j);
            xm = kb(i - 1, j);
            xp = kb(i + 1, j);
            ym = kb(i, j - 1);
            yp = kb(i, j + 1);
        } else {
            c = xm = xp = ym = yp = INT_MIN;
        }
    }
    // (bitwise & and | throughout: the predicates are cheap and evaluating all of them keeps the
    // sweep free of branches)
    __device__ __forceinline__ bool outside(int k) const { return bounded & ((k < 0) | (k >= Nz)); }
    __device__ __forceinline__ bool inactive(int kb, int k) const { return (k < kb) | outside(k); }
    // immersed_inactive_node of a cell centre, of the x or y face between two cells, and of the
    // z face below level k
    __device__ __forceinline__ bool iin(int kb, int k) const { return inactive(kb, k) & !outside(k); }
    __device__ __forceinline__ bool iin_face(int ka, int kb, int k) const {
        return inactive(ka, k) & inactive(kb, k) & !outside(k);
    }
    __device__ __forceinline__ bool iin_zface(int k) const {
        return inactive(c, k) & inactive(c, k - 1) & !(outside(k) & outside(k - 1));
    }
};

// V∇²ᶜᶜᶜ at level k of a column (conjugate_gradient_poisson_solver.jl:62-66) in FP64 from the seven
// stencil values.  On an immersed grid the inner ∂ vanishes across a face next to an immersed cell
// and the outer δ next to an immersed inactive face (conditional_differences.jl:15-44).
template <class T>
__device__ __forceinline__ double vlap_col(const DevGrid<T> &g, const ColumnMask<T> &m, int k, double c0, double xm, double xp,
                                           double ym, double yp, double zm, double zp) {
    const double idx = 1.0 / (double)g.dx, idy = 1.0 / (double)g.dy;
    const double Ax = (double)g.dy * (double)g.dzc[k], Ay = (double)g.dx * (double)g.dzc[k];
    const double Az = (double)g.dx * (double)g.dy;
    double gxp = (xp - c0) * idx, gxm = (c0 - xm) * idx;
    double gyp = (yp - c0) * idy, gym = (c0 - ym) * idy;
    // 1 / (double)Δzᶠ from the grid's FP64 table: the same correctly rounded quotient as dividing here
    double gzp = (zp - c0) * g.idzf64[k + 1], gzm = (c0 - zm) * g.idzf64[k];
    const bool ic = m.iin(m.c, k);
    gxp = (ic | m.iin(m.xp, k)) ? 0.0 : gxp;
    gxm = (ic | m.iin(m.xm, k)) ? 0.0 : gxm;
    gyp = (ic | m.iin(m.yp, k)) ? 0.0 : gyp;
    gym = (ic | m.iin(m.ym, k)) ? 0.0 : gym;
    gzp = (ic | m.iin(m.c, k + 1)) ? 0.0 : gzp;
    gzm = (ic | m.iin(m.c, k - 1)) ? 0.0 : gzm;
    const bool ox = m.iin_face(m.c, m.xm, k) | m.iin_face(m.xp, m.c, k);
    const bool oy = m.iin_face(m.c, m.ym, k) | m.iin_face(m.yp, m.c, k);
    const bool oz = m.iin_zface(k) | m.iin_zface(k + 1);
    const double tx = ox ? 0.0 : Ax * gxp - Ax * gxm;
    const double ty = oy ? 0.0 : Ay * gyp - Ay * gym;
    const double tz = oz ? 0.0 : Az * gzp - Az * gzm;
    return (tx + ty) + tz;
}

// ---------------------------------------------------------------------------------------------
// Sweeps of the 7-point operator.  A block takes a 64 × 4 tile of columns and a chunk of levels
// (cg_stencil_shape); each thread marches up its column with the values at k − 1, k and k + 1 in
// registers, so a level costs five loads (the level above and the four horizontal neighbours),
// and the immersed-boundary rules use the five bottom indices of ColumnMask.  Blocks are dealt to
// the XCDs in contiguous runs of work items (xcd_work_item), the chunks of one tile consecutive.
//   STENCIL_APPLY     q = V∇²p; partial sums of p·q, Σp and Σq
//   STENCIL_RESIDUAL  r = b − V∇²x, through the initial VolumeInverseNorm; partial sums of (r V⁻¹)²
//   STENCIL_TRUE      e = b − V∇²x; partial sums of e
//   STENCIL_ENERGY    partial sums of −d·V∇²d and, when gout is given, gout = −2·s·active·V∇²d
// ---------------------------------------------------------------------------------------------
enum { STENCIL_APPLY = 0, STENCIL_RESIDUAL = 1, STENCIL_TRUE = 2, STENCIL_ENERGY = 3 };

template <class T, int MODE>
__global__ void __launch_bounds__(kCgStTx * kCgStTy)
    k_cg_stencil(DevGrid<T> g, const double *__restrict__ src, const double *__restrict__ b, double *__restrict__ out,
                 float *__restrict__ gout, double gscale, double *__restrict__ partial, int kchunk, int ntx) {
    const int a = threadIdx.x & (kCgStTx - 1), c = threadIdx.x / kCgStTx;
    const int item = xcd_work_item((int)blockIdx.x, (int)gridDim.x);
    const int nch = (g.Nz + kchunk - 1) / kchunk;
    const int tile = item / nch, chunk = item - tile * nch;
    const int ty = tile / ntx, tx = tile - ty * ntx;
    const int i = tx * kCgStTx + a, j = ty * kCgStTy + c;
    const int k0 = chunk * kchunk, k1 = min(k0 + kchunk, g.Nz);
    double acc = 0, accp = 0, accq = 0;   // (STENCIL_APPLY: p·q, Σp, Σq)
    if (i < g.Nx && j < g.Ny) {
        const ColumnMask<T> m(g, i, j);
        const int64_t col = idx3(i, j, 0, g.sy, g.sz), sy = g.sy, sz = g.sz;
        const double *s = src + col;
        auto emit = [&](int k, double c0, double v) {
            const int64_t o = col + (int64_t)k * sz;
            if (MODE == STENCIL_APPLY) {
                out[o] = v;
                acc += c0 * v;
                accp += c0;
                accq += v;
            } else if (MODE == STENCIL_RESIDUAL) {
                const double rv = b[o] - v;
                const double rs = rv * (double)(T(1) / vol(g, k));
                acc += rs * rs;
                out[o] = rs * (double)vol(g, k);
            } else if (MODE == STENCIL_TRUE) {
                const double e = b[o] - v;
                out[o] = e;
                acc += e;
            } else {
                acc -= c0 * v;
                if (gout) {
                    const double act = m.inactive(m.c, k) ? 0.0 : 1.0;
                    gout[(int64_t)i + (int64_t)g.Nx * ((int64_t)j + (int64_t)g.Ny * k)] = (float)(-2.0 * gscale * act * v);
                }
            }
        };
        // the column's values at k − 1 and k carried in registers; each batch of kCgStBatch levels
        // loads its levels' four horizontal neighbours and the levels above (all loads issued first)
        double zm = s[(int64_t)(k0 - 1) * sz], c0 = s[(int64_t)k0 * sz];
        int k = k0;
        for (; k + kCgStBatch <= k1; k += kCgStBatch) {
            double xm[kCgStBatch], xp[kCgStBatch], ym[kCgStBatch], yp[kCgStBatch], zc[kCgStBatch];
#pragma unroll
            for (int u = 0; u < kCgStBatch; ++u) {
                const int64_t o = (int64_t)(k + u) * sz;
                xm[u] = s[o - 1]; xp[u] = s[o + 1]; ym[u] = s[o - sy]; yp[u] = s[o + sy];
                zc[u] = s[o + sz];   // level k + u + 1 ≤ k1 ≤ Nz: at most the top halo layer
            }
#pragma unroll
            for (int u = 0; u < kCgStBatch; ++u) {
                const double v = vlap_col(g, m, k + u, c0, xm[u], xp[u], ym[u], yp[u], zm, zc[u]);
                emit(k + u, c0, v);
                zm = c0;
                c0 = zc[u];
            }
        }
        for (; k < k1; ++k) {
            const int64_t o = (int64_t)k * sz;
            const double zp = s[o + sz];
            const double v = vlap_col(g, m, k, c0, s[o - 1], s[o + 1], s[o - sy], s[o + sy], zm, zp);
            emit(k, c0, v);
            zm = c0;
            c0 = zp;
        }
    }
    const double sum = block_sum(acc);
    if (MODE == STENCIL_APPLY) {
        const double sp = block_sum(accp), sq = block_sum(accq);
        if (threadIdx.x == 0) {
            partial[blockIdx.x] = sum;
            partial[gridDim.x + blockIdx.x] = sp;
            partial[2 * gridDim.x + blockIdx.x] = sq;
        }
    } else if (threadIdx.x == 0) {
        partial[blockIdx.x] = sum;
    }
}

// ---------------------------------------------------------------------------------------------
// mode 0: f ·= s, 1: f /= s, 2: f = s (interior)
template <class T> __global__ void __launch_bounds__(256) k_cg_scale(DevGrid<T> g, T *__restrict__ f, T s, int mode) {
    for_cells(g, [&](int, int, int, int64_t o, int64_t) { f[o] = mode == 2 ? s : (mode == 1 ? f[o] / s : f[o] * s); });
}

template <class T> __global__ void __launch_bounds__(256) k_cg_gather(DevGrid<T> g, const double *__restrict__ x, T *__restrict__ dense) {
    for_cells(g, [&](int, int, int, int64_t o, int64_t t) { dense[t] = (T)x[o]; });
}

// x = (double)(f · s) (the warm start, solve_for_pressure.jl:150-153), or f = (T)x
template <class T> __global__ void __launch_bounds__(256) k_cg_xio(DevGrid<T> g, T *__restrict__ f, double *__restrict__ x, T s, int store) {
    for_cells(g, [&](int, int, int, int64_t o, int64_t) {
        if (store) f[o] = (T)x[o];
        else x[o] = (double)(f[o] * s);
    });
}

// b = active · divᶜᶜᶜ(U) · V (solve_for_pressure.jl:44-50), or b = rhs · V for a given dense rhs
// (solve!(x, ::ConjugateGradientPoissonSolver, b) takes the caller's right-hand side as it is)
template <class T> __global__ void __launch_bounds__(256) k_cg_source(Ctx<T> C, const T *__restrict__ dense_rhs, double *__restrict__ b) {
    const DevGrid<T> &g = C.g;
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        T d;
        if (dense_rhs) d = dense_rhs[t];
        else d = C.divU(i, j, k, i + 1, j + 1);
        const T act = (!dense_rhs && IB<T>(g).inactive_cell(i, j, k)) ? T(0) : T(1);
        b[o] = (double)((act * d) * vol(g, k));
    });
}

// DiagonallyDominantPreconditioner (conjugate_gradient_poisson_solver.jl:256-296); r has its halos
template <class T> __device__ __forceinline__ double dd_ac(const DevGrid<T> &g, int k) {
    const double ax = (double)(g.dy * g.dzc[k]) * (double)(T(1) / g.dx);
    const double ay = (double)(g.dx * g.dzc[k]) * (double)(T(1) / g.dy);
    const double azm = (double)g.Az * (double)(T(1) / g.dzf[k]);
    const double azp = (double)g.Az * (double)(T(1) / g.dzf[k + 1]);
    return ((((-ax - ax) - ay) - ay) - azm) - azp;
}
template <class T> __global__ void __launch_bounds__(256) k_cg_dd(DevGrid<T> g, const double *__restrict__ r, T *__restrict__ z) {
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        const double ax = (double)(g.dy * g.dzc[k]) * (double)(T(1) / g.dx);
        const double ay = (double)(g.dx * g.dzc[k]) * (double)(T(1) / g.dy);
        const double azm = (double)g.Az * (double)(T(1) / g.dzf[k]);
        const double azp = (double)g.Az * (double)(T(1) / g.dzf[k + 1]);
        const double ac = dd_ac(g, k), acm = dd_ac(g, k - 1), acp = dd_ac(g, k + 1);
        double acc = r[o];
        acc = acc - ax / ac * r[o - 1];
        acc = acc - ax / ac * r[o + 1];
        acc = acc - ay / ac * r[o - g.sy];
        acc = acc - ay / ac * r[o + g.sy];
        acc = acc - azm / acm * r[o - g.sz];
        acc = acc - azp / acp * r[o + g.sz];
        z[o] = (T)((1.0 / fabs(ac) * acc) * active(g, i, j, k));
    });
}

// ColumnwiseTridiagonalPreconditioner (conjugate_gradient_poisson_solver.jl:298-397): one thread per
// column solves the vertical part of V∇² (diagonal × (1 + 1/100), identity rows for inactive and
// vertically isolated cells) by the Thomas sweep of batched_tridiagonal_solver.jl in the model's
// precision; a pivot under 10·eps keeps z's previous value, as the reference's forward sweep does.
// The sweep's ratios t live in a dense scratch array indexed k·(Nx·Ny) + column (coalesced).
template <class T> __global__ void __launch_bounds__(256) k_cg_ct(DevGrid<T> g, const double *r, T *z, T *t) {
    const int64_t ncol = (int64_t)g.Nx * g.Ny;
    const int64_t col = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    const int i = (int)(col % g.Nx), j = (int)(col / g.Nx);
    const IB<T> ib(g);
    auto azf = [&](int k) { return g.Az * g.idzf[k]; };
    auto off = [&](int k) {
        const T ap = ib.inactive_cell(i, j, k + 1) ? T(0) : azf(k + 1);
        return ib.inactive_cell(i, j, k) ? T(0) : ap;
    };
    auto diag = [&](int k) {
        const bool s = ib.inactive_cell(i, j, k), b = ib.inactive_cell(i, j, k - 1), a = ib.inactive_cell(i, j, k + 1);
        const T am = b ? T(0) : azf(k), ap = a ? T(0) : azf(k + 1);
        const T eps = T(1) / T(100);
        return (s || (b && a)) ? T(1) : -(am + ap) * (T(1) + eps);
    };
    const T guard = T(10) * std::numeric_limits<T>::epsilon();
    const int64_t o = idx3(i, j, 0, g.sy, g.sz);
    T beta = diag(0);
    T zp = T(r[o]) / beta;
    z[o] = zp;
    for (int k = 1; k < g.Nz; ++k) {
        const T c = off(k - 1), bk = diag(k);
        const T tk = c / beta;
        t[(int64_t)k * ncol + col] = tk;
        beta = bk - c * tk;
        const int64_t ok = o + (int64_t)k * g.sz;
        const T phis = (T(r[ok]) - c * zp) / beta;
        const T zk = fabs(beta) > guard ? phis : z[ok];
        z[ok] = zk;
        zp = zk;
    }
    for (int k = g.Nz - 2; k >= 0; --k) {
        const int64_t ok = o + (int64_t)k * g.sz;
        const T zk = z[ok] - t[(int64_t)(k + 1) * ncol + col] * zp;
        z[ok] = zk;
        zp = zk;
    }
}

template <class T> __global__ void __launch_bounds__(256) k_cg_copy_r(DevGrid<T> g, const double *__restrict__ r, T *__restrict__ z) {
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        z[o] = (T)r[o];
    });
}

// partial sums of z·r
template <class T> __global__ void __launch_bounds__(256) k_cg_dot_zr(DevGrid<T> g, const T *__restrict__ z, const double *__restrict__ r, double *partial) {
    double acc = 0;
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        acc += (double)z[o] * r[o];
    });
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// p = z (first iteration, it = 0) or p = z + β p with β = ρ / ρ_prev, or the flexible form
// β = (ρ − z·r_prev) / ρ_prev, from the iteration's scalars on the device (cg_slot_rho); `it` is the
// iteration index or any number of its parity that is nonzero after the first iteration.  The
// cells next to the domain's faces also write p's one-layer halos (the fill of cg_halo_r: periodic
// x and y, z by its topology), which the operator sweep that follows reads.
template <class T>
__global__ void __launch_bounds__(256) k_cg_update_p(DevGrid<T> g, double *__restrict__ p, const T *__restrict__ z,
                                                     const double *__restrict__ sc, int it, int flex) {
    const int first = it == 0;
    double beta = 0.0;
    if (!first) {
        const double rho = sc[cg_slot_rho(it)], prev = sc[cg_slot_rho(it - 1)];
        beta = flex ? (rho - sc[cg_slot_rho(it) + 1]) / prev : rho / prev;
    }
    const int64_t xs = g.Nx, ys = (int64_t)g.Ny * g.sy, zs = (int64_t)g.Nz * g.sz;
    const bool zper = g.tz == OM_PERIODIC;
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        const double pv = first ? (double)z[o] : (double)z[o] + beta * p[o];
        p[o] = pv;
        if (i == 0) p[o + xs] = pv;
        if (i == g.Nx - 1) p[o - xs] = pv;
        if (j == 0) p[o + ys] = pv;
        if (j == g.Ny - 1) p[o - ys] = pv;
        if (k == 0) p[zper ? o + zs : o - g.sz] = pv;
        if (k == g.Nz - 1) p[zper ? o - zs : o + g.sz] = pv;
    });
}


// The gauge and norm sweep (conjugate_gradient_poisson_solver.jl:87-96, then the VolumeInverseNorm
// of :13-34): x ← (x − mean x)·active, r ← (r − mean r)·active, partial sums of (r V⁻¹)², and of the
// new x and r (the next iteration's means start from them).  With `dense`, the FFT preconditioner's
// source r V⁻¹ of the next iteration is written as well.
//   step = false: the solve's first sweep, means 0 (no update);
//   step = true:  the iteration's update x += α p, r −= α q comes first, α = ρ / p·Ap, and the means
//                 of the updated x and r come from sums taken before it: Σx + α Σp and Σr − α Σq
//                 (Σx, Σr from the previous sweep, Σp, Σq from the operator sweep).  This is the
//                 same mean the reference takes after the update, summed in another order, so
//                 the update and the gauge share one pass over x, p, r and q.
template <class T>
__global__ void __launch_bounds__(256) k_cg_gauge_norm(DevGrid<T> g, double *__restrict__ x, const double *__restrict__ p,
                                                       double *__restrict__ r, const double *__restrict__ q,
                                                       const double *__restrict__ sc, int it, int step, double ncell,
                                                       T *__restrict__ dense, double *partial) {
    double alpha = 0.0, mx = 0.0, mr = 0.0;
    if (step) {
        alpha = sc[cg_slot_rho(it)] / sc[kCgSlotPQ];
        mx = (sc[kCgSlotGauge + 1] + alpha * sc[kCgSlotPQ + 1]) / ncell;
        mr = (sc[kCgSlotGauge + 2] - alpha * sc[kCgSlotPQ + 2]) / ncell;
    }
    double acc = 0, sx = 0, sr = 0;
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        const double act = active(g, i, j, k);
        const double xu = step ? x[o] + alpha * p[o] : x[o];
        const double ru = step ? r[o] - alpha * q[o] : r[o];
        const double xg = (xu - mx) * act;
        x[o] = xg;
        const double rg = (ru - mr) * act;
        const double rs = rg * (double)(T(1) / vol(g, k));
        acc += rs * rs;
        const double rn = rs * (double)vol(g, k);
        r[o] = rn;
        sx += xg;
        sr += rn;
        if (dense) dense[t] = (T)(rn * (double)(T(1) / vol(g, k)));   // r V⁻¹ of the stored r (conjugate_gradient_poisson_solver.jl:169-211)
    });
    const double s0 = block_sum(acc), s1 = block_sum(sx), s2 = block_sum(sr);
    if (threadIdx.x == 0) { partial[blockIdx.x] = s0; partial[kRedBlocks + blockIdx.x] = s1; partial[2 * kRedBlocks + blockIdx.x] = s2; }
}


// ‖V⁻¹(e − ē)‖² partial sums: the recomputed residual in the zero-mean gauge the iteration uses
// (the mean of b − A x is the part of b outside the operator's range, which no x removes)
template <class T> __global__ void __launch_bounds__(256) k_cg_gauged_norm(DevGrid<T> g, const double *e, double mean,
                                                                           double *partial) {
    double acc = 0;
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        const double v = (e[o] - mean) / (double)vol(g, k);
        acc += v * v;
    });
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------------------------
// Learned correction to the FFT preconditioner (the network is in om_learned.hip) and flexible CG
// ---------------------------------------------------------------------------------------------
// partial sums over active cells of (r V⁻¹)², z² and 1 (the input scales)
template <class T> __global__ void __launch_bounds__(256) k_ln_scales(DevGrid<T> g, const double *__restrict__ r, const T *__restrict__ z, double *partial) {
    double a0 = 0, a1 = 0, a2 = 0;
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        const double act = active(g, i, j, k);
        const double rv = r[o] * (double)(T(1) / vol(g, k)), zv = (double)z[o];
        a0 += act * rv * rv;
        a1 += act * zv * zv;
        a2 += act;
    });
    const double s0 = block_sum(a0), s1 = block_sum(a1), s2 = block_sum(a2);
    if (threadIdx.x == 0) { partial[blockIdx.x] = s0; partial[kRedBlocks + blockIdx.x] = s1; partial[2 * kRedBlocks + blockIdx.x] = s2; }
}
// network input channels, dense FP32 in cell order: r V⁻¹ / s_r, z / s_z, active
template <class T> __global__ void __launch_bounds__(256) k_ln_input(DevGrid<T> g, const double *__restrict__ r, const T *__restrict__ z, double isr,
                                                                     double isz, float *__restrict__ in3) {
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        in3[3 * t] = (float)(r[o] * (double)(T(1) / vol(g, k)) * isr);
        in3[3 * t + 1] = (float)((double)z[o] * isz);
        in3[3 * t + 2] = (float)active(g, i, j, k);
    });
}
// z += s_z · active · (network output)
// the same three channels written straight into the network's padded BF16 input box (16 channels
// per voxel, 3..15 zero; halos filled afterwards by the network's halo kernel): each channel is the
// FP32 value k_ln_input stores, rounded to BF16 as the dense input's conversion rounds it
template <class T> __global__ void __launch_bounds__(256) k_ln_x0(DevGrid<T> g, const double *__restrict__ r, const T *__restrict__ z,
                                                                  double isr, double isz, __bf16 *__restrict__ x0) {
    typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        bf16x8 a, b;
        for (int e = 0; e < 8; ++e) { a[e] = (__bf16)0.0f; b[e] = (__bf16)0.0f; }
        a[0] = (__bf16)(float)(r[o] * (double)(T(1) / vol(g, k)) * isr);
        a[1] = (__bf16)(float)((double)z[o] * isz);
        a[2] = (__bf16)(float)active(g, i, j, k);
        const int64_t v = ((int64_t)(k + 1) * (g.Ny + 2) + (j + 1)) * (int64_t)(g.Nx + 2) + (i + 1);
        __bf16 *p = x0 + v * 16;
        *reinterpret_cast<bf16x8 *>(p) = a;
        *reinterpret_cast<bf16x8 *>(p + 8) = b;
    });
}
// z += s_z · m · N: a non-finite network output (an overflowed activation) adds nothing, so the
// preconditioned residual stays finite and the solve's stall test can hand over to the FFT
// preconditioner
template <class T> __global__ void __launch_bounds__(256) k_ln_correct(DevGrid<T> g, T *__restrict__ z, const float *__restrict__ out, double sz) {
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        const float y = out[t];
        if (isfinite(y)) z[o] = (T)((double)z[o] + sz * active(g, i, j, k) * (double)y);
    });
}
// interior copy of an FP64 cell field
template <class T> __global__ void __launch_bounds__(256) k_cg_copy_d(DevGrid<T> g, const double *__restrict__ src, double *__restrict__ dst) {
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        dst[o] = src[o];
    });
}
// z·r and z·r_prev (the flexible-CG β of Polak–Ribière form)
template <class T> __global__ void __launch_bounds__(256) k_cg_dot2(DevGrid<T> g, const T *__restrict__ z, const double *__restrict__ r, const double *__restrict__ rp,
                                                                    double *partial) {
    double a0 = 0, a1 = 0;
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        const double zv = (double)z[o];
        a0 += zv * r[o];
        a1 += zv * rp[o];
    });
    const double s0 = block_sum(a0), s1 = block_sum(a1);
    if (threadIdx.x == 0) { partial[blockIdx.x] = s0; partial[kRedBlocks + blockIdx.x] = s1; }
}
// target of a logged pair: (x_final − x_k) / s_z − ẑ_k on active cells (0 elsewhere), with partial
// sums of the target and of the active count (its mean is removed afterwards: the pressure's
// constant is free)
template <class T> __global__ void __launch_bounds__(256) k_ln_target(DevGrid<T> g, const double *xf, const double *xk, const float *in3k,
                                                                      double isz, float *tgt, double *partial) {
    double a0 = 0, a1 = 0;
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        const double act = active(g, i, j, k);
        const double v = act * ((xf[o] - xk[o]) * isz - (double)in3k[3 * t + 1]);
        tgt[t] = (float)v;
        a0 += v;
        a1 += act;
    });
    const double s0 = block_sum(a0), s1 = block_sum(a1);
    if (threadIdx.x == 0) { partial[blockIdx.x] = s0; partial[kRedBlocks + blockIdx.x] = s1; }
}
template <class T> __global__ void __launch_bounds__(256) k_ln_center(DevGrid<T> g, float *tgt, float mean) {
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        tgt[t] = active(g, i, j, k) > 0 ? tgt[t] - mean : 0.0f;
    });
}
template <class T> __global__ void __launch_bounds__(256) k_ln_mask(DevGrid<T> g, float *mask) {
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        mask[t] = (float)active(g, i, j, k);
    });
}

// A-energy training loss of the learned correction.  d = act · (y − t) (dimensionless, in units
// of rms(z_FFT)) goes into a padded FP64 field; after its halos, the loss is dᵀ(−V∇²)d with the
// solver's own (conditional) operator, and its gradient with respect to y is −2·act·V∇²d.
template <class T> __global__ void __launch_bounds__(256) k_ln_err(DevGrid<T> g, const float *y, const float *tgt, double *d) {
    for_cells(g, [&](int i, int j, int k, int64_t o, int64_t t) {
        d[idx3(i, j, k, g.sy, g.sz)] = active(g, i, j, k) * ((y ? (double)y[t] : 0.0) - (double)tgt[t]);
    });
}

// out[v] = Σ partial[v·nper + b] over b < nper, in a fixed order
static __global__ void __launch_bounds__(kRedThreads) k_cg_finish(const double *partial, int nvals, int nper, double *out) {
    for (int v = 0; v < nvals; ++v) {
        double acc = 0;
        for (int b = threadIdx.x; b < nper; b += kRedThreads) acc += partial[v * nper + b];
        const double s = block_sum(acc);
        if (threadIdx.x == 0) out[v] = s;
    }
}

// halos of an FP64 cell field for the 7-point operators: periodic wrap in x and y (one layer),
// Bounded z one-layer copy (NoFlux), Periodic z one-layer wrap
template <class T> __global__ void __launch_bounds__(256) k_cg_halo_r(DevGrid<T> g, double *r, int pass) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (pass == 0) {   // z faces over the interior x-y extent
        const int i = a % g.Nx, j = a / g.Nx;
        if (j >= g.Ny) return;
        const int64_t c = idx3(i, j, 0, g.sy, g.sz);
        if (g.tz == OM_PERIODIC) {
            r[c - g.sz] = r[c + (int64_t)(g.Nz - 1) * g.sz];
            r[c + (int64_t)g.Nz * g.sz] = r[c];
        } else {
            r[c - g.sz] = r[c];
            r[c + (int64_t)g.Nz * g.sz] = r[c + (int64_t)(g.Nz - 1) * g.sz];
        }
    } else if (pass == 1) {   // x halos over y, z interior
        const int j = a % g.Ny, k = a / g.Ny;
        if (k >= g.Nz) return;
        const int64_t c = idx3(0, j, k, g.sy, g.sz);
        r[c - 1] = r[c + g.Nx - 1];
        r[c + g.Nx] = r[c];
    } else {   // y halos over x, z interior
        const int i = a % g.Nx, k = a / g.Nx;
        if (k >= g.Nz) return;
        const int64_t c = idx3(i, 0, k, g.sy, g.sz);
        r[c - g.sy] = r[c + (int64_t)(g.Ny - 1) * g.sy];
        r[c + (int64_t)g.Ny * g.sy] = r[c];
    }
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
static dim3 red_grid() { return dim3(kRedBlocks); }
static dim3 red_blk() { return dim3(kRedThreads); }

template <class T, int MODE>
static void cg_stencil(const DevGrid<T> &g, const double *src, const double *b, double *out, float *gout, double gscale,
                       double *partial, double *outv, hipStream_t st) {
    const CgStencilShape s = cg_stencil_shape(g.Nx, g.Ny, g.Nz);
    hipLaunchKernelGGL((k_cg_stencil<T, MODE>), dim3(s.nblocks), dim3(kCgStTx * kCgStTy), 0, st, g, src, b, out, gout, gscale,
                       partial, s.kchunk, s.ntx);
    hipLaunchKernelGGL(k_cg_finish, dim3(1), red_blk(), 0, st, (const double *)partial, MODE == STENCIL_APPLY ? 3 : 1, s.nblocks,
                       outv);
}

template <class T> void cg_scale(const DevGrid<T> &g, T *f, T s, int mode, hipStream_t st) {
    hipLaunchKernelGGL(k_cg_scale<T>, red_grid(), red_blk(), 0, st, g, f, s, mode);
}
template <class T> void cg_gather(const DevGrid<T> &g, const double *x, T *dense, hipStream_t st) {
    hipLaunchKernelGGL(k_cg_gather<T>, red_grid(), red_blk(), 0, st, g, x, dense);
}
template <class T> void cg_xio(const DevGrid<T> &g, T *f, double *x, T s, int store, hipStream_t st) {
    hipLaunchKernelGGL(k_cg_xio<T>, red_grid(), red_blk(), 0, st, g, f, x, s, store);
}
template <class T> void cg_source(const Ctx<T> &C, const T *dense_rhs, double *b, hipStream_t st) {
    hipLaunchKernelGGL(k_cg_source<T>, red_grid(), red_blk(), 0, st, C, dense_rhs, b);
}
template <class T> void cg_residual(const DevGrid<T> &g, const double *x, const double *b, double *r, double *partial, double *out,
                                    hipStream_t st) {
    cg_stencil<T, STENCIL_RESIDUAL>(g, x, b, r, (float *)nullptr, 0.0, partial, out, st);
}
template <class T> void cg_halo_r(const DevGrid<T> &g, double *r, hipStream_t st) {
    const int nxy = g.Nx * g.Ny, nyz = g.Ny * g.Nz, nxz = g.Nx * g.Nz;
    if (g.tz != OM_FLAT) hipLaunchKernelGGL(k_cg_halo_r<T>, dim3((nxy + 255) / 256), dim3(256), 0, st, g, r, 0);
    hipLaunchKernelGGL(k_cg_halo_r<T>, dim3((nyz + 255) / 256), dim3(256), 0, st, g, r, 1);
    hipLaunchKernelGGL(k_cg_halo_r<T>, dim3((nxz + 255) / 256), dim3(256), 0, st, g, r, 2);
}
template <class T> void cg_dd(const DevGrid<T> &g, const double *r, T *z, hipStream_t st) {
    hipLaunchKernelGGL(k_cg_dd<T>, red_grid(), red_blk(), 0, st, g, r, z);
}
template <class T> void cg_ct(const DevGrid<T> &g, const double *r, T *z, T *t, hipStream_t st) {
    const int64_t ncol = (int64_t)g.Nx * g.Ny;
    hipLaunchKernelGGL(k_cg_ct<T>, dim3((unsigned)((ncol + 255) / 256)), dim3(256), 0, st, g, r, z, t);
}
template <class T> void ln_scales(const DevGrid<T> &g, const double *r, const T *z, double *partial, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_ln_scales<T>, red_grid(), red_blk(), 0, st, g, r, z, partial);
    hipLaunchKernelGGL(k_cg_finish, dim3(1), red_blk(), 0, st, (const double *)partial, 3, kRedBlocks, out);
}
template <class T> void ln_input(const DevGrid<T> &g, const double *r, const T *z, double isr, double isz, float *in3, hipStream_t st) {
    hipLaunchKernelGGL(k_ln_input<T>, red_grid(), red_blk(), 0, st, g, r, z, isr, isz, in3);
}
template <class T> void ln_x0(const DevGrid<T> &g, const double *r, const T *z, double isr, double isz, void *x0, hipStream_t st) {
    hipLaunchKernelGGL(k_ln_x0<T>, red_grid(), red_blk(), 0, st, g, r, z, isr, isz, (__bf16 *)x0);
}
template <class T> void ln_correct(const DevGrid<T> &g, T *z, const float *out, double sz, hipStream_t st) {
    hipLaunchKernelGGL(k_ln_correct<T>, red_grid(), red_blk(), 0, st, g, z, out, sz);
}
template <class T> void cg_copy_d(const DevGrid<T> &g, const double *src, double *dst, hipStream_t st) {
    hipLaunchKernelGGL(k_cg_copy_d<T>, red_grid(), red_blk(), 0, st, g, src, dst);
}
template <class T> void cg_dot2(const DevGrid<T> &g, const T *z, const double *r, const double *rp, double *partial, double *out,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_cg_dot2<T>, red_grid(), red_blk(), 0, st, g, z, r, rp, partial);
    hipLaunchKernelGGL(k_cg_finish, dim3(1), red_blk(), 0, st, (const double *)partial, 2, kRedBlocks, out);
}
template <class T> void ln_target(const DevGrid<T> &g, const double *xf, const double *xk, const float *in3k, double isz, float *tgt,
                                  double *partial, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_ln_target<T>, red_grid(), red_blk(), 0, st, g, xf, xk, in3k, isz, tgt, partial);
    hipLaunchKernelGGL(k_cg_finish, dim3(1), red_blk(), 0, st, (const double *)partial, 2, kRedBlocks, out);
}
template <class T> void ln_center(const DevGrid<T> &g, float *tgt, float mean, hipStream_t st) {
    hipLaunchKernelGGL(k_ln_center<T>, red_grid(), red_blk(), 0, st, g, tgt, mean);
}
template <class T> void ln_energy(const DevGrid<T> &g, const float *y, const float *tgt, double *dscratch, double gscale, float *gout,
                                  double *partial, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_ln_err<T>, red_grid(), red_blk(), 0, st, g, y, tgt, dscratch);
    cg_halo_r(g, dscratch, st);
    cg_stencil<T, STENCIL_ENERGY>(g, dscratch, (const double *)nullptr, (double *)nullptr, gout, gscale, partial, out, st);
}
template <class T> void ln_mask(const DevGrid<T> &g, float *mask, hipStream_t st) {
    hipLaunchKernelGGL(k_ln_mask<T>, red_grid(), red_blk(), 0, st, g, mask);
}
template <class T> void cg_copy_r(const DevGrid<T> &g, const double *r, T *z, hipStream_t st) {
    hipLaunchKernelGGL(k_cg_copy_r<T>, red_grid(), red_blk(), 0, st, g, r, z);
}
template <class T> void cg_dot_zr(const DevGrid<T> &g, const T *z, const double *r, double *partial, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_cg_dot_zr<T>, red_grid(), red_blk(), 0, st, g, z, r, partial);
    hipLaunchKernelGGL(k_cg_finish, dim3(1), red_blk(), 0, st, (const double *)partial, 1, kRedBlocks, out);
}
template <class T> void cg_update_p(const DevGrid<T> &g, double *p, const T *z, const double *sc, int it, int flex, hipStream_t st) {
    hipLaunchKernelGGL(k_cg_update_p<T>, red_grid(), red_blk(), 0, st, g, p, z, sc, it, flex);
}
template <class T> void cg_apply(const DevGrid<T> &g, const double *p, double *q, double *partial, double *out, hipStream_t st) {
    cg_stencil<T, STENCIL_APPLY>(g, p, (const double *)nullptr, q, (float *)nullptr, 0.0, partial, out, st);
}
template <class T> void cg_gauge_norm(const DevGrid<T> &g, double *x, const double *p, double *r, const double *q,
                                      const double *sc, int it, int step, double ncell, T *dense, double *partial, double *out,
                                      hipStream_t st) {
    hipLaunchKernelGGL(k_cg_gauge_norm<T>, red_grid(), red_blk(), 0, st, g, x, p, r, q, sc, it, step, ncell, dense, partial);
    hipLaunchKernelGGL(k_cg_finish, dim3(1), red_blk(), 0, st, (const double *)partial, 3, kRedBlocks, out);
}
template <class T> void cg_true_residual(const DevGrid<T> &g, const double *x, const double *b, double *e, double *partial,
                                         double *out, hipStream_t st) {
    cg_stencil<T, STENCIL_TRUE>(g, x, b, e, (float *)nullptr, 0.0, partial, out, st);
}
template <class T> void cg_gauged_norm(const DevGrid<T> &g, const double *e, double mean, double *partial, double *out,
                                       hipStream_t st) {
    hipLaunchKernelGGL(k_cg_gauged_norm<T>, red_grid(), red_blk(), 0, st, g, e, mean, partial);
    hipLaunchKernelGGL(k_cg_finish, dim3(1), red_blk(), 0, st, (const double *)partial, 1, kRedBlocks, out);
}

#define OM_CG_INSTANTIATE(T)                                                                                     \
    template void cg_scale<T>(const DevGrid<T> &, T *, T, int, hipStream_t);                                    \
    template void cg_gather<T>(const DevGrid<T> &, const double *, T *, hipStream_t);                           \
    template void cg_xio<T>(const DevGrid<T> &, T *, double *, T, int, hipStream_t);                            \
    template void cg_source<T>(const Ctx<T> &, const T *, double *, hipStream_t);                               \
    template void cg_residual<T>(const DevGrid<T> &, const double *, const double *, double *, double *, double *, \
                                 hipStream_t);                                                                   \
    template void cg_halo_r<T>(const DevGrid<T> &, double *, hipStream_t);                                      \
    template void cg_dd<T>(const DevGrid<T> &, const double *, T *, hipStream_t);                               \
    template void cg_copy_r<T>(const DevGrid<T> &, const double *, T *, hipStream_t);                           \
    template void cg_ct<T>(const DevGrid<T> &, const double *, T *, T *, hipStream_t);                          \
    template void ln_scales<T>(const DevGrid<T> &, const double *, const T *, double *, double *, hipStream_t);    \
    template void ln_input<T>(const DevGrid<T> &, const double *, const T *, double, double, float *, hipStream_t); \
    template void ln_x0<T>(const DevGrid<T> &, const double *, const T *, double, double, void *, hipStream_t);     \
    template void ln_correct<T>(const DevGrid<T> &, T *, const float *, double, hipStream_t);                    \
    template void cg_copy_d<T>(const DevGrid<T> &, const double *, double *, hipStream_t);                       \
    template void cg_dot2<T>(const DevGrid<T> &, const T *, const double *, const double *, double *, double *, hipStream_t); \
    template void ln_target<T>(const DevGrid<T> &, const double *, const double *, const float *, double, float *, double *, \
                               double *, hipStream_t);                                                            \
    template void ln_center<T>(const DevGrid<T> &, float *, float, hipStream_t);                                 \
    template void ln_mask<T>(const DevGrid<T> &, float *, hipStream_t);                                          \
    template void ln_energy<T>(const DevGrid<T> &, const float *, const float *, double *, double, float *, double *, double *, \
                               hipStream_t);                                                                      \
    template void cg_dot_zr<T>(const DevGrid<T> &, const T *, const double *, double *, double *, hipStream_t);  \
    template void cg_update_p<T>(const DevGrid<T> &, double *, const T *, const double *, int, int, hipStream_t); \
    template void cg_apply<T>(const DevGrid<T> &, const double *, double *, double *, double *, hipStream_t);   \
    template void cg_gauge_norm<T>(const DevGrid<T> &, double *, const double *, double *, const double *,      \
                                   const double *, int, int, double, T *, double *, double *, hipStream_t);     \
    template void cg_true_residual<T>(const DevGrid<T> &, const double *, const double *, double *, double *,   \
                                      double *, hipStream_t);                                                    \
    template void cg_gauged_norm<T>(const DevGrid<T> &, const double *, double, double *, double *, hipStream_t);

#ifndef OM_ONLY_DOUBLE
OM_CG_INSTANTIATE(float)
#endif
#ifndef OM_ONLY_FLOAT
OM_CG_INSTANTIATE(double)
#endif

}  // namespace om
