// om_fft.hip — pressure Poisson solvers on gfx950.
//
// Spectral array S (complex, x fastest): S[(k·Ny + j)·Ms + kx], kx < Mx = Nx/2 + 1 (real-to-complex
// along a Periodic x) or 1 (Flat x); the row stride Ms rounds Mx up to a 128-byte line, and the
// passes skip the padding entries kx ≥ Mx.  Passes, each one HBM read + one write of S:
//   X-forward  : source term (∇·U* formed on the fly, optionally ×Δzᶜ) → real-to-complex FFT
//   Y-forward  : complex FFT along y                    (skipped when y is the last axis)
//   last axis  : forward transform · eigenvalue division · backward transform in one kernel
//                (Bounded z: DCT-II / DCT-III pair through one complex FFT of length Nz, or the
//                 2Nz-point even extension for odd Nz; Fourier–tridiagonal: Thomas algorithm)
//   Y-backward, X-backward (complex-to-real, written into the pressure field's interior)
// Each line transform is an in-place radix-4 Stockham FFT in LDS for power-of-two lengths and a
// direct DFT otherwise; the global loads that stage the lines are batched per thread.  Reference: src/Solvers/fft_based_poisson_solver.jl:94-124,
// src/Solvers/fourier_tridiagonal_poisson_solver.jl:231-260, src/Solvers/batched_tridiagonal_solver.jl:211-243.
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "om_internal.h"
#include "om_batch.h"
#include "om_kernels.h"

namespace om {

template <class T> using C2 = typename Vec2<T>::type;

template <class T> __device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { C2<T> r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <class T> __device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { C2<T> r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <class T> __device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) { C2<T> r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r; }
template <class T> __device__ __forceinline__ C2<T> cconj(C2<T> a) { a.y = -a.y; return a; }
template <class T> __device__ __forceinline__ C2<T> cmk(T x, T y) { C2<T> r; r.x = x; r.y = y; return r; }

// Global → LDS staging.  get(t) issues the global loads of element t, put(t, v) stores it in LDS.
// Each thread first issues the loads of B elements and only then stores them, so a wave keeps B
// loads in flight (a plain load-then-store loop waits out the full memory latency per element).
template <int B, class Get, class Put>
__device__ __forceinline__ void staged_copy(int count, Get get, Put put) {
    const int nt = blockDim.x;
    for (int t0 = threadIdx.x; t0 < count; t0 += B * nt) {
        decltype(get(0)) v[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int t = t0 + u * nt;
            if (t < count) v[u] = get(t);
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int t = t0 + u * nt;
            if (t < count) put(t, v[u]);
        }
    }
}

// R lines of length n, row r at a + r·ld; tw (in LDS) holds exp(-2πi m / n), m < n.
// Power-of-two n: in-place Stockham stages of radix 4 (one radix-2 stage first when log2 n is
// odd).  A radix-R stage at sub-length m maps a[kk + jj·m + p·n/R] (p < R) to
// a[kk + R·jj·m + q·m] = w^{q·jj·m} Σ_p c_p e^{∓2πi pq/R}; every thread reads its butterflies'
// inputs into registers, the block synchronises, and the outputs overwrite the same buffer, so
// one buffer serves every stage.  Each thread holds at most 2 radix-4 (4 radix-2) butterflies:
// R·n ≤ 8·blockDim.x.  Index splits are shifts and masks.
// Other n: a direct DFT into b.  Returns the buffer that holds the (unnormalised) transform.
template <class T>
__device__ C2<T> *lds_fft(C2<T> *a, C2<T> *b, int R, int n, int ld, const C2<T> *tw, bool inverse) {
    const int nt = blockDim.x;
    const int tid = threadIdx.x;
    if ((n & (n - 1)) == 0) {
        const int lg = __builtin_ctz((unsigned)n);
        int lm = 0;   // log2 of the current sub-length m
        if (lg & 1) {
            const int half = n >> 1, lh = lg - 1, nb = R * half;
            C2<T> c0[4], c1[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = tid + u * nt;
                if (t < nb) {
                    const int base = (t >> lh) * ld + (t & (half - 1));
                    c0[u] = a[base];
                    c1[u] = a[base + half];
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = tid + u * nt;
                if (t < nb) {
                    const int jj = t & (half - 1);
                    C2<T> w = tw[jj];
                    if (inverse) w.y = -w.y;
                    const int ob = (t >> lh) * ld + (jj << 1);
                    a[ob] = cadd<T>(c0[u], c1[u]);
                    a[ob + 1] = cmul<T>(w, csub<T>(c0[u], c1[u]));
                }
            }
            __syncthreads();
            lm = 1;
        }
        const int quarter = n >> 2, lq = lg - 2, nb = R * quarter;
        for (; lm < lg; lm += 2) {
            const int m = 1 << lm;
            C2<T> c[2][4];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int t = tid + u * nt;
                if (t < nb) {
                    const int ib = (t >> lq) * ld + (t & (quarter - 1));   // = row + kk + jj·m
#pragma unroll
                    for (int p = 0; p < 4; ++p) c[u][p] = a[ib + p * quarter];
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int t = tid + u * nt;
                if (t < nb) {
                    const int bf = t & (quarter - 1);
                    const int jj = bf >> lm, kk = bf & (m - 1);
                    const C2<T> t0 = cadd<T>(c[u][0], c[u][2]), t1 = csub<T>(c[u][0], c[u][2]);
                    const C2<T> t2 = cadd<T>(c[u][1], c[u][3]), d = csub<T>(c[u][1], c[u][3]);
                    const C2<T> t3 = inverse ? cmk<T>(-d.y, d.x) : cmk<T>(d.y, -d.x);   // ±i·(c1 − c3)
                    const int e = jj << lm;
                    C2<T> w1 = tw[e], w2 = tw[2 * e], w3 = tw[3 * e];
                    if (inverse) { w1.y = -w1.y; w2.y = -w2.y; w3.y = -w3.y; }
                    const int ob = (t >> lq) * ld + kk + (jj << (lm + 2));
                    a[ob] = cadd<T>(t0, t2);
                    a[ob + m] = cmul<T>(w1, cadd<T>(t1, t3));
                    a[ob + 2 * m] = cmul<T>(w2, csub<T>(t0, t2));
                    a[ob + 3 * m] = cmul<T>(w3, csub<T>(t1, t3));
                }
            }
            __syncthreads();
        }
        return a;
    }
    for (int t = tid; t < R * n; t += nt) {
        const int r = t / n, ko = t - r * n;
        C2<T> s = cmk<T>(T(0), T(0));
        for (int q = 0; q < n; ++q) {
            C2<T> w = tw[(int)(((int64_t)ko * q) % n)];
            if (inverse) w.y = -w.y;
            s = cadd<T>(s, cmul<T>(a[r * ld + q], w));
        }
        b[r * ld + ko] = s;
    }
    __syncthreads();
    return b;
}

template <class T> __device__ __forceinline__ void load_table(C2<T> *dst, const C2<T> *src, int n) {
    staged_copy<4>(n, [&](int t) { return src[t]; }, [&](int t, C2<T> v) { dst[t] = v; });
}

constexpr int kFftThreads = 512;   // threads per block of the line kernels (a multiple of R ≤ 16)

// Line staging geometry, chosen on the host: R lines per block, row stride ld (n + 1 when R > 1,
// so the R rows start in different LDS banks), tables first, then buffer a, then b (only for a
// length that is not a power of two).
struct LineGeom {
    int R, lgR, ld;
};

// ---------------------------------------------------------------------------------------------
// X-forward: source term → R2C along x.  One block handles R rows (j, k).
// src == nullptr: the source is ∇·U of the state (solve_for_pressure.jl:12-18); else a dense
// interior array.  scale_dzc: multiply by Δzᶜ(k) (Fourier–tridiagonal source, :36-42).
// ---------------------------------------------------------------------------------------------
template <class T> struct XFwdArgs {
    Ctx<T> C;
    const T *src;
    C2<T> *S;
    const C2<T> *tw;      // length Nx/2 table (for the half-length complex FFT), Nx when Nx is odd
    const C2<T> *post;    // exp(-2πi k / Nx), k = 0 .. Nx/2
    int scale_dzc;
    LineGeom L;
    FastDiv fNy, fMx;
    int Ms;
};

template <class T> __device__ __forceinline__ T source_at(const Ctx<T> &C, const T *src, int i, int j, int k) {
    if (src) return src[(int64_t)i + (int64_t)C.g.Nx * ((int64_t)j + (int64_t)C.g.Ny * k)];
    // divergence with periodic index wrap in x and y (halos of U* are not needed)
    const int ip = (i + 1 == C.g.Nx) ? 0 : i + 1;
    const int jp = (j + 1 == C.g.Ny) ? 0 : j + 1;
    if (C.g.kbot) return C.divU(i, j, k, ip, jp);
    // without an immersed bottom: Ctx::divU's expression, the six loads addressed from one index
    const int64_t o = C.at(i, j, k);
    const T *u = C.s.U[0] + o, *v = C.s.U[1] + o, *w = C.s.U[2] + o;
    const T sx = C.flx() ? T(0) : C.Axc(k) * u[ip - i] - C.Axc(k) * u[0];
    const T sy = C.fly() ? T(0) : C.Ayc(k) * v[(int64_t)(jp - j) * C.g.sy] - C.Ayc(k) * v[0];
    const T sz = C.flz() ? T(0) : C.g.Az * w[C.g.sz] - C.g.Az * w[0];
    return C.iVc(k) * ((sx + sy) + sz);
}

template <class T> struct KXFwd {
    static constexpr int kThreads = kFftThreads, kMinBlocks = 1;
    static __device__ __forceinline__ void run(const XFwdArgs<T> &A, const Blk &B);
};
template <class T> __device__ __forceinline__ void KXFwd<T>::run(const XFwdArgs<T> &A, const Blk &B) {
    extern __shared__ unsigned char smem_raw[];
    const Ctx<T> &C = A.C;
    const int Nx = C.g.Nx, Ny = C.g.Ny;
    const bool half = (Nx & 1) == 0;
    const int M = half ? Nx / 2 : Nx;      // line length of the complex FFT
    const int Mx = Nx / 2 + 1;
    const int R = A.L.R, ld = A.L.ld;
    C2<T> *twl = reinterpret_cast<C2<T> *>(smem_raw);
    C2<T> *a = twl + M;
    C2<T> *b = a + R * ld;
    const int row0 = B.x * R;
    const int nrows = Ny * C.g.Nz;
    load_table<T>(twl, A.tw, M);
    T *ap = reinterpret_cast<T *>(a);
    // without an immersed bottom the source is ∇·U of three non-Flat axes on a whole row: each
    // thread forms its (at most kFastX) elements from loads all issued before any arithmetic,
    // with no per-element branch (the general path below tests the topology per element).  Its
    // neighbours across the x and y edges are read by wrapped index, and a Bounded z's boundary
    // faces of w as the zero the impenetrable fill writes there, so it reads no halo
    // (poisson_source_reads_no_halo is the same test on the host)
    constexpr int kFastX = kXFwdFastElems;
    const bool fast = !A.src && !C.g.kbot && !C.flx() && !C.fly() && !C.flz() && half && !A.scale_dzc &&
                      Nx <= kFastX * (int)(blockDim.x >> A.L.lgR);
    if (fast) {
        const int lgt = __builtin_ctz((unsigned)blockDim.x) - A.L.lgR;   // log2 of threads per row
        const int tpr = 1 << lgt;
        const int r = threadIdx.x >> lgt, l = threadIdx.x & (tpr - 1);
        const int row = row0 + r;
        const bool live = row < nrows;
        const int k = live ? (int)fdiv((uint32_t)row, A.fNy) : 0, j = live ? row - k * Ny : 0;
        const int jp = (j + 1 == Ny) ? 0 : j + 1;
        const int64_t o = C.at(0, j, k);
        const T *u = C.s.U[0] + o, *v = C.s.U[1] + o, *w = C.s.U[2] + o;
        const T *vp = C.s.U[1] + C.at(0, jp, k), *wp = w + C.g.sz;
        const bool wb = C.g.tz == OM_BOUNDED && k == 0, wt = C.g.tz == OM_BOUNDED && k == C.g.Nz - 1;
        T su[kFastX], sq[kFastX], sv[kFastX], svp[kFastX], sw[kFastX], swp[kFastX];
#pragma unroll
        for (int e = 0; e < kFastX; ++e) {
            const int i = l + e * tpr;
            const bool in = live && i < Nx;
            const int ii = in ? i : 0;
            const int ip = (ii + 1 == Nx) ? 0 : ii + 1;
            su[e] = in ? u[ii] : T(0);
            sq[e] = in ? u[ip] : T(0);
            sv[e] = in ? v[ii] : T(0);
            svp[e] = in ? vp[ii] : T(0);
            sw[e] = in && !wb ? w[ii] : T(0);
            swp[e] = in && !wt ? wp[ii] : T(0);
        }
        const T Ax = C.Axc(k), Ay = C.Ayc(k), Az = C.g.Az, iV = C.iVc(k);
#pragma unroll
        for (int e = 0; e < kFastX; ++e) {
            const int i = l + e * tpr;
            if (i < Nx) {
                // Ctx::divU's expression (source_at), in its order of operations
                const T sx = Ax * sq[e] - Ax * su[e];
                const T sy = Ay * svp[e] - Ay * sv[e];
                const T sz = Az * swp[e] - Az * sw[e];
                ap[2 * (r * ld + (i >> 1)) + (i & 1)] = iV * ((sx + sy) + sz);
            }
        }
    } else {
        // blockDim.x / R threads per row: a thread's row (j, k), its metric factors and addresses are
        // fixed, and its elements i = l, l + tpr, … differ by constant offsets; the loads of four
        // elements are issued before their LDS stores.  Rows past the last stage zeros.
        const int lgt = __builtin_ctz((unsigned)blockDim.x) - A.L.lgR;   // log2 of threads per row
        const int tpr = 1 << lgt;
        const int r = threadIdx.x >> lgt, l = threadIdx.x & (tpr - 1);
        const int row = row0 + r;
        const bool live = row < nrows;
        const int k = live ? (int)fdiv((uint32_t)row, A.fNy) : 0, j = live ? row - k * Ny : 0;
        const T zs = A.scale_dzc && live ? C.g.dzc[k] : T(1);
        constexpr int NB = 4;
        for (int i0 = l; i0 < Nx; i0 += NB * tpr) {
            T x[NB];
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const int i = i0 + u * tpr;
                x[u] = T(0);
                if (live && i < Nx) {
                    x[u] = source_at<T>(C, A.src, i, j, k);
                    if (A.scale_dzc) x[u] = zs * x[u];
                }
            }
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const int i = i0 + u * tpr;
                if (i < Nx) {
                    if (half) ap[2 * (r * ld + (i >> 1)) + (i & 1)] = x[u];
                    else a[r * ld + i] = cmk<T>(x[u], T(0));
                }
            }
        }
    }
    __syncthreads();
    const C2<T> *Z = lds_fft<T>(a, b, R, M, ld, twl, false);
    for (int t = threadIdx.x; t < R * Mx; t += blockDim.x) {
        const int r = (int)fdiv((uint32_t)t, A.fMx), kk = t - r * Mx;
        const int row = row0 + r;
        if (row >= nrows) continue;
        if (!half) { A.S[(int64_t)row * A.Ms + kk] = Z[r * ld + kk]; continue; }
        const C2<T> zk = Z[r * ld + (kk == M ? 0 : kk)];
        const C2<T> zc = cconj<T>(Z[r * ld + (kk == 0 ? 0 : M - kk)]);
        const C2<T> E = cmk<T>(T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y + zc.y));
        // O = (zk - zc) / (2i) = (-i/2)(zk - zc)
        const C2<T> d = csub<T>(zk, zc);
        const C2<T> O = cmk<T>(T(0.5) * d.y, T(-0.5) * d.x);
        A.S[(int64_t)row * A.Ms + kk] = cadd<T>(E, cmul<T>(A.post[kk], O));
    }
}

// ---------------------------------------------------------------------------------------------
// X-backward: C2R along x, writes out[i, j, k] = x / n_norm (+ optional mean subtraction)
// ---------------------------------------------------------------------------------------------
template <class T> struct XBwdArgs {
    DevGrid<T> g;
    const C2<T> *S;
    const C2<T> *tw, *post;
    T *out;            // field interior view (strided) or dense array
    int dense;
    T scale;           // 1 / (Nx · Ny · Nz-normalisations)
    T divisor;         // the written value is (x · scale) / divisor
    LineGeom L;
    FastDiv fNy;
    int Ms;
};

template <class T> struct KXBwd {
    static constexpr int kThreads = kFftThreads, kMinBlocks = 1;
    static __device__ __forceinline__ void run(const XBwdArgs<T> &A, const Blk &B);
};
template <class T> __device__ __forceinline__ void KXBwd<T>::run(const XBwdArgs<T> &A, const Blk &B) {
    extern __shared__ unsigned char smem_raw[];
    const int Nx = A.g.Nx, Ny = A.g.Ny;
    const bool half = (Nx & 1) == 0;
    const int M = half ? Nx / 2 : Nx;
    const int Mx = Nx / 2 + 1;
    const int R = A.L.R, ld = A.L.ld;
    C2<T> *twl = reinterpret_cast<C2<T> *>(smem_raw);
    C2<T> *a = twl + M;
    C2<T> *b = a + R * ld;
    const int row0 = B.x * R;
    const int nrows = Ny * A.g.Nz;
    load_table<T>(twl, A.tw, M);
    // blockDim.x / R threads per row, as in k_xfwd: a thread's row and its addresses are fixed
    const int lgt = __builtin_ctz((unsigned)blockDim.x) - A.L.lgR;   // log2 of threads per row
    const int tpr = 1 << lgt;
    const int r = threadIdx.x >> lgt, l = threadIdx.x & (tpr - 1);
    const int row = row0 + r;
    const bool live = row < nrows;
    {
        const C2<T> *Srow = A.S + (int64_t)(live ? row : 0) * A.Ms;
        constexpr int B = 4;
        for (int k0 = l; k0 < M; k0 += B * tpr) {
            C2<T> v[B];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const int kk = k0 + u * tpr;
                C2<T> Zk = cmk<T>(T(0), T(0));
                if (live && kk < M) {
                    if (!half) {
                        Zk = kk < Mx ? Srow[kk] : cconj<T>(Srow[Nx - kk]);
                    } else {
                        const C2<T> Xk = Srow[kk];
                        const C2<T> Xc = cconj<T>(Srow[M - kk]);
                        const C2<T> E = cmk<T>(T(0.5) * (Xk.x + Xc.x), T(0.5) * (Xk.y + Xc.y));
                        const C2<T> d = csub<T>(Xk, Xc);
                        // O = (Xk - conj(X_{M-k})) · e^{+2πik/N} / 2 ;  Z = E + i O
                        const C2<T> O = cmul<T>(cmk<T>(T(0.5) * d.x, T(0.5) * d.y), cconj<T>(A.post[kk]));
                        Zk = cmk<T>(E.x - O.y, E.y + O.x);
                    }
                }
                v[u] = Zk;
            }
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const int kk = k0 + u * tpr;
                if (kk < M) a[r * ld + kk] = v[u];
            }
        }
    }
    __syncthreads();
    const C2<T> *z = lds_fft<T>(a, b, R, M, ld, twl, true);
    if (live) {
        const int k = (int)fdiv((uint32_t)row, A.fNy), j = row - k * Ny;
        T *dst = A.dense ? A.out + (int64_t)Nx * ((int64_t)j + (int64_t)Ny * k) : A.out + idx3(0, j, k, A.g.sy, A.g.sz);
        for (int i = l; i < Nx; i += tpr) {
            T x;
            if (half) { const C2<T> zz = z[r * ld + (i >> 1)]; x = (i & 1) ? zz.y : zz.x; }
            else x = z[r * ld + i].x;
            dst[i] = ft_div<T>(x * A.scale, A.divisor);   // IEEE in FP64; FP32 by the reciprocal of the divisor
        }
    }
}

// −v / den, and 0 for the zero mode
template <class T> __device__ __forceinline__ C2<T> eig_divide(C2<T> v, T den, bool zero_mode) {
    return zero_mode ? cmk<T>(T(0), T(0)) : cmk<T>(ft_div<T>(-v.x, den), ft_div<T>(-v.y, den));
}

// ---------------------------------------------------------------------------------------------
// Y pass: complex FFT along y.  Lines ℓ = k·Ms + kx (ℓ < Ms·Nz, Ms the padded row stride), R
// consecutive lines per block; element j of line ℓ sits at S[(k·Ny + j)·Ms + kx].  Lines with
// kx ≥ Mx are row padding and are skipped.
// fused_divide: forward → divide by eigenvalues → backward (y is the last non-Flat axis).
// ---------------------------------------------------------------------------------------------
template <class T> struct YArgs {
    DevGrid<T> g;
    C2<T> *S;
    const C2<T> *tw;
    int Mx, Ms;
    int inverse;
    int fused_divide;
    const T *lx, *ly;   // eigenvalues (fused)
    LineGeom L;
    FastDiv fMs;        // division by the row stride Ms
};

template <class T> struct KYPass {
    static constexpr int kThreads = kFftThreads, kMinBlocks = 1;
    static __device__ __forceinline__ void run(const YArgs<T> &A, const Blk &B);
};
template <class T> __device__ __forceinline__ void KYPass<T>::run(const YArgs<T> &A, const Blk &B) {
    extern __shared__ unsigned char smem_raw[];
    const int Ny = A.g.Ny, Mx = A.Mx, R = A.L.R, lgR = A.L.lgR, ld = A.L.ld;
    C2<T> *twl = reinterpret_cast<C2<T> *>(smem_raw);
    C2<T> *a = twl + Ny;
    C2<T> *b = a + R * ld;
    const int Ms = A.Ms;
    const int nlines = Ms * A.g.Nz;
    // blockDim.x is a multiple of R, so every element a thread stages belongs to the same line
    const int r = threadIdx.x & (R - 1);
    const int line = B.x * R + r;
    const int k = (int)fdiv((uint32_t)line, A.fMs), kx = line - k * Ms;
    const bool valid = line < nlines && kx < Mx;                         // padding lines: nothing
    C2<T> *col = A.S + (int64_t)k * Ny * Ms + kx;
    load_table<T>(twl, A.tw, Ny);
    staged_copy<8>(R * Ny,
        [&](int t) { return valid ? col[(int64_t)(t >> lgR) * Ms] : cmk<T>(T(0), T(0)); },
        [&](int t, C2<T> v) { a[r * ld + (t >> lgR)] = v; });
    __syncthreads();
    C2<T> *res = lds_fft<T>(a, b, R, Ny, ld, twl, A.inverse != 0);
    if (A.fused_divide) {
        C2<T> *other = (res == a) ? b : a;
        for (int t = threadIdx.x; t < R * Ny; t += blockDim.x) {
            const int j = t >> lgR;
            if (valid) res[r * ld + j] = eig_divide<T>(res[r * ld + j], A.lx[kx] + A.ly[j], kx == 0 && j == 0);
        }
        __syncthreads();
        res = lds_fft<T>(res, other, R, Ny, ld, twl, true);
    }
    if (!valid) return;
    for (int t = threadIdx.x; t < R * Ny; t += blockDim.x) {
        const int j = t >> lgR;
        col[(int64_t)j * Ms] = res[r * ld + j];
    }
}

// ---------------------------------------------------------------------------------------------
// Z pass (last axis): forward transform, eigenvalue division, backward transform.  Lines
// ℓ = j·Ms + kx (ℓ < Ms·Ny), R consecutive lines per block; element k of line ℓ sits at
// S[k·Ny·Ms + ℓ], so a block reads R contiguous values per level; lines with kx ≥ Mx are row
// padding and are skipped.
//   zmode 0 (Periodic z): FFT of length Nz.
//   zmode 1 (Bounded z, Nz even): the DCT-II of length N = Nz through one complex FFT of length N
//     (Makhoul 1980): v = (x0, x2, …, x_{N−2}, x_{N−1}, …, x3, x1), V = FFT(v),
//     X_k = c_k V_k + c̄_k V_{N−k} with c_k = e^{−iπk/2N}; after the division the inverse is
//     V'_k = c̄_k (Φ_k − i Φ_{N−k}) (Φ_N = 0), v' = IFFT(V'), x read back through the same
//     permutation.  Each thread handles the pair (k, N−k).  The result equals the inverse
//     transform of the 2N-point even extension, so it shares that normalisation (1 / 2N).
//   zmode 2 (Bounded z, Nz odd): FFT of the 2N-point even extension, mirrored in LDS.
// ---------------------------------------------------------------------------------------------
template <class T> struct ZArgs {
    DevGrid<T> g;
    C2<T> *S;
    const C2<T> *tw;         // table for the FFT length L
    const C2<T> *dct;        // c_k, k < Nz (zmode 1)
    int Mx, Ms, L, zmode;    // L = Nz (zmodes 0, 1) or 2Nz (zmode 2); Ms: row stride of S
    const T *lx, *ly, *lz;   // lz: Nz entries (Periodic), 2Nz (Bounded, symmetric)
    LineGeom G;
};

template <class T> struct KZPass {
    static constexpr int kThreads = kFftThreads, kMinBlocks = 1;
    static __device__ __forceinline__ void run(const ZArgs<T> &A, const Blk &B);
};
template <class T> __device__ __forceinline__ void KZPass<T>::run(const ZArgs<T> &A, const Blk &B) {
    extern __shared__ unsigned char smem_raw[];
    const int Ny = A.g.Ny, Nz = A.g.Nz, Mx = A.Mx, L = A.L, zmode = A.zmode;
    const int R = A.G.R, lgR = A.G.lgR, ld = A.G.ld;
    C2<T> *twl = reinterpret_cast<C2<T> *>(smem_raw);
    C2<T> *dctl = twl + L;
    C2<T> *a = dctl + (zmode == 1 ? Nz : 0);
    C2<T> *b = a + R * ld;
    const int Ms = A.Ms;
    const int nlines = Ms * Ny;
    const int r = threadIdx.x & (R - 1);   // blockDim.x is a multiple of R
    const int line = B.x * R + r;
    const int j = line < nlines ? line / Ms : 0, kx = line - j * Ms;
    const bool valid = line < nlines && kx < Mx;   // padding lines: nothing
    const int64_t plane = (int64_t)Ny * Ms;
    C2<T> *col = A.S + line;
    load_table<T>(twl, A.tw, L);
    if (zmode == 1) load_table<T>(dctl, A.dct, Nz);
    staged_copy<8>(R * Nz,
        [&](int t) { return valid ? col[(int64_t)(t >> lgR) * plane] : cmk<T>(T(0), T(0)); },
        [&](int t, C2<T> v) {
            const int k = t >> lgR;
            if (zmode == 1) a[r * ld + ((k & 1) ? Nz - ((k + 1) >> 1) : (k >> 1))] = v;
            else {
                a[r * ld + k] = v;
                if (zmode == 2) a[r * ld + (L - 1 - k)] = v;
            }
        });
    __syncthreads();
    C2<T> *res = lds_fft<T>(a, b, R, L, ld, twl, false);
    C2<T> *other = (res == a) ? b : a;
    const T lxy = valid ? A.lx[kx] + A.ly[j] : T(1);
    const bool zero_col = kx == 0 && j == 0;
    if (zmode == 1) {
        const int hN = Nz >> 1;
        for (int t = threadIdx.x; t < R * (hN + 1); t += blockDim.x) {
            const int kk = t >> lgR, p = kk == 0 ? 0 : Nz - kk;
            C2<T> *row = res + r * ld;
            const C2<T> Vk = row[kk], Vp = row[p];
            const C2<T> ck = dctl[kk], cp = dctl[p];
            const C2<T> Xk = cadd<T>(cmul<T>(ck, Vk), cmul<T>(cconj<T>(ck), Vp));
            const C2<T> Xp = cadd<T>(cmul<T>(cp, Vp), cmul<T>(cconj<T>(cp), Vk));
            const C2<T> Fk = eig_divide<T>(Xk, lxy + A.lz[kk], zero_col && kk == 0);
            const C2<T> Fp = kk == 0 ? cmk<T>(T(0), T(0)) : eig_divide<T>(Xp, lxy + A.lz[p], false);
            // V'_k = c̄_k (Φ_k − i Φ_{N−k}),  −i·z = (z.y, −z.x)
            row[kk] = cmul<T>(cconj<T>(ck), cmk<T>(Fk.x + Fp.y, Fk.y - Fp.x));
            if (p != kk) row[p] = cmul<T>(cconj<T>(cp), cmk<T>(Fp.x + Fk.y, Fp.y - Fk.x));
        }
    } else {
        for (int t = threadIdx.x; t < R * L; t += blockDim.x) {
            const int k = t >> lgR;
            res[r * ld + k] = eig_divide<T>(res[r * ld + k], lxy + A.lz[k], zero_col && k == 0);
        }
    }
    __syncthreads();
    res = lds_fft<T>(res, other, R, L, ld, twl, true);
    if (!valid) return;
    for (int t = threadIdx.x; t < R * Nz; t += blockDim.x) {
        const int k = t >> lgR;
        const int pos = zmode == 1 ? ((k & 1) ? Nz - ((k + 1) >> 1) : (k >> 1)) : k;
        col[(int64_t)k * plane] = res[r * ld + pos];
    }
}

// ---------------------------------------------------------------------------------------------
// Fourier–tridiagonal z solve (Thomas per (kx, ky) column; diagonals from
// fourier_tridiagonal_poisson_solver.jl:199-229).  One thread per column, the columns ℓ = ky·Ms + kx
// flattened so that a level of a block's columns is one contiguous run (level k of column ℓ at
// k·(Ms·Ny) + ℓ; columns with kx ≥ Mx are row padding and are skipped); t[k] kept in LDS.  Each sweep
// reads its levels in batches of tri_batch (16 FP32, 8 FP64), all
// loads of a batch issued before its arithmetic, so a column costs 2·Nz/batch memory round
// trips instead of 2·Nz.  The (0,0) column is singular: its last pivot is a rounding residue and
// is elided as in the reference; the constant this leaves is removed with the column mean
// (ϕ -= mean(ϕ), :253).
// ---------------------------------------------------------------------------------------------
template <class T> struct TriArgs {
    DevGrid<T> g;
    C2<T> *S;
    int Mx, Ms;
    const T *lx, *ly;
    T eps10;
    T zscale;   // (1 / L_backward_norm) applied later; unused here
};

template <class T> constexpr int tri_batch() { return sizeof(T) == 4 ? 16 : 8; }

// The level coefficients and the spectrum come in as restrict-qualified arguments: the sweeps'
// stores to S then provably leave them alone, so their (column-uniform) loads stay scalar.
template <class T>
__global__ void __launch_bounds__(64) k_tridiag(TriArgs<T> A, C2<T> *__restrict__ Sg, const T *__restrict__ dzc,
                                                const T *__restrict__ idzf) {
    extern __shared__ unsigned char smem_raw[];
    T *tt = reinterpret_cast<T *>(smem_raw);   // [Nz][blockDim.x]
    constexpr int KB = tri_batch<T>();
    const int Ny = A.g.Ny, Nz = A.g.Nz, Mx = A.Mx;
    const int bw = blockDim.x;
    const int64_t ncol = (int64_t)A.Ms * Ny;   // columns ℓ = ky·Ms + kx, padding columns kx ≥ Mx skipped
    const int64_t col = (int64_t)blockIdx.x * bw + threadIdx.x;
    if (col >= ncol) return;
    const int j = (int)(col / A.Ms), kx = (int)(col - (int64_t)j * A.Ms);
    if (kx >= Mx) return;
    const T lam = A.lx[kx] + A.ly[j];
    auto diag = [&](int k) -> T {   // 0-based k
        const T dz = dzc[k];
        if (Nz == 1) return -dz * lam;
        if (k == 0) return -idzf[1] - dz * lam;
        if (k == Nz - 1) return -idzf[Nz - 1] - dz * lam;
        return -(idzf[k + 1] + idzf[k]) - dz * lam;
    };
    auto off = [&](int k) -> T { return idzf[k + 1]; };   // a[k] = c[k] = 1/Δzᶠ(k+1) (reference q = k+1)
    C2<T> *S = Sg + col;
    auto at = [&](int k) -> int64_t { return (int64_t)k * ncol; };
    // forward sweep: level 0, then full batches of KB levels (loads first), then the rest
    // (ft_div: IEEE division in FP64, as the reference's Thomas sweep; the hardware reciprocal in
    // FP32, which shortens the serial chain through β at each level)
    T beta = diag(0);
    C2<T> phi;
    {
        const C2<T> f = S[0];
        phi = cmk<T>(ft_div<T>(f.x, beta), ft_div<T>(f.y, beta));
        S[0] = phi;
    }
    auto fwd = [&](int k, const C2<T> &f) {
        const T c = off(k - 1), a = off(k - 1);
        const T t = ft_div<T>(c, beta);
        tt[k * bw + threadIdx.x] = t;
        beta = diag(k) - a * t;
        if (fabs(beta) > A.eps10) phi = cmk<T>(ft_div<T>(f.x - a * phi.x, beta), ft_div<T>(f.y - a * phi.y, beta));
        else phi = cmk<T>(T(0), T(0));
        S[at(k)] = phi;
    };
    int k = 1;
    for (; k + KB <= Nz; k += KB) {
        C2<T> f[KB];
#pragma unroll
        for (int u = 0; u < KB; ++u) f[u] = S[at(k + u)];
#pragma unroll
        for (int u = 0; u < KB; ++u) fwd(k + u, f[u]);
    }
    for (; k < Nz; ++k) fwd(k, S[at(k)]);
    // backward substitution from level Nz − 2 down, in the same batches; the column's sum is
    // accumulated on the way (used by the (0,0) column only, top level first)
    C2<T> colsum = phi;
    auto bwd = [&](int kk, const C2<T> &p) {
        const T t = tt[(kk + 1) * bw + threadIdx.x];
        const C2<T> q = cmk<T>(p.x - t * phi.x, p.y - t * phi.y);
        S[at(kk)] = q;
        phi = q;
        colsum = cadd<T>(colsum, q);
    };
    k = Nz - 2;
    for (; k - KB + 1 >= 0; k -= KB) {
        C2<T> p[KB];
#pragma unroll
        for (int u = 0; u < KB; ++u) p[u] = S[at(k - u)];
#pragma unroll
        for (int u = 0; u < KB; ++u) bwd(k - u, p[u]);
    }
    for (; k >= 0; --k) bwd(k, S[at(k)]);
    if (col == 0) {
        // remove the mean of the (0,0) column: equals removing the volume mean of ϕ (the reference
        // subtracts the mean of the real field; the sum here runs down the column as the backward
        // sweep produced it).  The column is one thread's work at the end of the kernel, so its
        // read-modify-write pass goes in batches too (one round trip per level would set the
        // kernel's duration).
        const C2<T> s = cmk<T>(colsum.x / T(Nz), colsum.y / T(Nz));
        int q = 0;
        for (; q + KB <= Nz; q += KB) {
            C2<T> v[KB];
#pragma unroll
            for (int u = 0; u < KB; ++u) v[u] = S[at(q + u)];
#pragma unroll
            for (int u = 0; u < KB; ++u) S[at(q + u)] = csub<T>(v[u], s);
        }
        for (; q < Nz; ++q) S[at(q)] = csub<T>(S[at(q)], s);
    }
}

// ---------------------------------------------------------------------------------------------
// FP64 iterative refinement: residual of a computed solution and the correction update
// ---------------------------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) k_poisson_residual(Ctx<T> C, RefineOp L, const T *p, int64_t psy, int64_t psz,
                                                          const double *b64, double idiv, T *r) {
    const DevGrid<T> &g = C.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.Nx || j >= g.Ny) return;
    const int64_t dense = (int64_t)i + (int64_t)g.Nx * ((int64_t)j + (int64_t)g.Ny * k);
    const int ip = i + 1 == g.Nx ? 0 : i + 1, im = i == 0 ? g.Nx - 1 : i - 1;
    const int jp = j + 1 == g.Ny ? 0 : j + 1, jm = j == 0 ? g.Ny - 1 : j - 1;
    double b;
    if (b64) b = b64[dense];
    else {
        // ∇·U in FP64 from the stored velocities (divergence_operators.jl:16-19 on a rectilinear grid)
        b = 0;
        if (!C.flx()) b += ((double)C.ld(C.s.U[0], ip, j, k) - (double)C.ld(C.s.U[0], i, j, k)) * (double)g.idx;
        if (!C.fly()) b += ((double)C.ld(C.s.U[1], i, jp, k) - (double)C.ld(C.s.U[1], i, j, k)) * (double)g.idy;
        if (!C.flz())
            b += ((double)C.ld(C.s.U[2], i, j, k + 1) - (double)C.ld(C.s.U[2], i, j, k)) * (L.ihz2 < 0 ? L.idzc[k] : (double)g.idzc[k]);
        b *= idiv;
    }
    auto P = [&](int ii, int jj, int kk) -> double { return (double)p[(int64_t)ii + psy * jj + psz * kk]; };
    const double pc = P(i, j, k);
    double lap = 0;
    if (!C.flx()) lap += ((P(ip, j, k) - pc) - (pc - P(im, j, k))) * L.ihx2;
    if (!C.fly()) lap += ((P(i, jp, k) - pc) - (pc - P(i, jm, k))) * L.ihy2;
    if (!C.flz()) {
        const bool per = g.tz == OM_PERIODIC;
        const int kp = k + 1 == g.Nz ? (per ? 0 : -1) : k + 1, km = k == 0 ? (per ? g.Nz - 1 : -1) : k - 1;
        if (L.ihz2 >= 0) {
            const double up = kp >= 0 ? P(i, j, kp) - pc : 0.0, dn = km >= 0 ? pc - P(i, j, km) : 0.0;
            lap += (up - dn) * L.ihz2;
        } else {   // stretched: (flux through face k+1 − flux through face k) / Δzᶜ
            const double up = kp >= 0 ? (P(i, j, kp) - pc) * L.idzf[k + 1] : 0.0;
            const double dn = km >= 0 ? (pc - P(i, j, km)) * L.idzf[k] : 0.0;
            lap += (up - dn) * L.idzc[k];
        }
    }
    r[dense] = (T)(b - lap);
}

template <class T>
__global__ void __launch_bounds__(256) k_add_correction(DevGrid<T> g, T *p, int64_t psy, int64_t psz, const T *d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.Nx || j >= g.Ny) return;
    p[(int64_t)i + psy * j + psz * k] += d[(int64_t)i + (int64_t)g.Nx * ((int64_t)j + (int64_t)g.Ny * k)];
}

template <class T> void launch_poisson_residual(const Ctx<T> &C, const RefineOp &L, const T *p, int64_t psy, int64_t psz,
                                                const double *b64, double divisor, T *r, hipStream_t st) {
    const dim3 blk(64, 4, 1), grid((C.g.Nx + 63) / 64, (C.g.Ny + 3) / 4, C.g.Nz);
    hipLaunchKernelGGL(k_poisson_residual<T>, grid, blk, 0, st, C, L, p, psy, psz, b64, 1.0 / divisor, r);
}
template <class T> void launch_add_correction(const DevGrid<T> &g, T *p, int64_t psy, int64_t psz, const T *d, hipStream_t st) {
    const dim3 blk(64, 4, 1), grid((g.Nx + 63) / 64, (g.Ny + 3) / 4, g.Nz);
    hipLaunchKernelGGL(k_add_correction<T>, grid, blk, 0, st, g, p, psy, psz, d);
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
static int ilog2(int v) { int l = 0; while ((1 << l) < v) ++l; return l; }

// R lines of length n with `tables` complex table entries in front (om_kernels.h chooses R)
static LineGeom line_geom(int n, int tables, size_t e, size_t &lds) {
    const int R = fft_lines_per_block(n, tables, e);
    if (R < 1) throw std::runtime_error("FFT line does not fit in LDS");
    const int ld = R > 1 ? n + 1 : n;
    const bool p2 = (n & (n - 1)) == 0;
    lds = ((size_t)(p2 ? 1 : 2) * R * ld + tables) * e;
    return LineGeom{R, ilog2(R), ld};
}

template <class T> void launch_poisson(const PoissonPlan<T> &P, const Ctx<T> &C, const T *src, T *out, int dense_out,
                                       T divisor, hipStream_t st) {
    const DevGrid<T> &g = C.g;
    const size_t e = sizeof(C2<T>);
    const size_t budget = 64 * 1024;
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const int M = (Nx & 1) == 0 ? Nx / 2 : Nx;
    const int Mx = Nx / 2 + 1;
    const int nrows = Ny * Nz;
    size_t lds = 0;
    {
        const LineGeom L = line_geom(M, M, e, lds);
        XFwdArgs<T> A{C, src, P.S, P.tw_xh, P.post, P.tridiagonal, L, make_fastdiv((uint32_t)Ny), make_fastdiv((uint32_t)Mx), P.Ms};
        launch_k<KXFwd<T>>(dim3((nrows + L.R - 1) / L.R), dim3(kFftThreads), lds, st, A);
    }
    const bool z_active = g.tz != OM_FLAT;
    const LineGeom Ly = line_geom(Ny, Ny, e, lds);
    const size_t lds_y = lds;
    const int ygrid = (P.Ms * Nz + Ly.R - 1) / Ly.R;
    {
        YArgs<T> A{g, P.S, P.tw_y, Mx, P.Ms, 0, z_active ? 0 : 1, P.lx, P.ly, Ly, make_fastdiv((uint32_t)P.Ms)};
        launch_k<KYPass<T>>(dim3(ygrid), dim3(kFftThreads), lds_y, st, A);
    }
    T scale = T(1) / (T(M) * T(Ny));
    if (z_active) {
        if (P.tridiagonal) {
            TriArgs<T> A{g, P.S, Mx, P.Ms, P.lx, P.ly, P.eps10, T(1)};
            // columns per block (one wave, partly filled when narrow): each column is a serial sweep,
            // so the launch wants waves on every SIMD (≥ 2048 of them) more than full waves; 16
            // columns still read whole 128-byte lines of the spectrum.  Limited by the LDS holding t[k].
            const int64_t ncol = (int64_t)P.Ms * Ny;
            int bw = 64;
            while (bw > 16 && ncol / bw < 2048) bw >>= 1;
            while (bw > 1 && (size_t)Nz * bw * sizeof(T) > budget) bw >>= 1;
            hipLaunchKernelGGL(k_tridiag<T>, dim3((unsigned)((ncol + bw - 1) / bw)), dim3(bw), (size_t)Nz * bw * sizeof(T), st, A,
                               P.S, g.dzc, g.idzf);
        } else {
            const LineGeom Lz = line_geom(P.Lz_line, P.Lz_line + (P.zmode == 1 ? Nz : 0), e, lds);
            ZArgs<T> A{g, P.S, P.tw_z, P.tw_dct, Mx, P.Ms, P.Lz_line, P.zmode, P.lx, P.ly, P.lz, Lz};
            launch_k<KZPass<T>>(dim3((P.Ms * Ny + Lz.R - 1) / Lz.R), dim3(kFftThreads), lds, st, A);
            scale = T(1) / ((T(M) * T(Ny)) * T(P.Lz_norm));
        }
        YArgs<T> A{g, P.S, P.tw_y, Mx, P.Ms, 1, 0, P.lx, P.ly, Ly, make_fastdiv((uint32_t)P.Ms)};
        launch_k<KYPass<T>>(dim3(ygrid), dim3(kFftThreads), lds_y, st, A);
    }
    {
        const LineGeom L = line_geom(M, M, e, lds);
        XBwdArgs<T> A{g, P.S, P.tw_xh, P.post, out, dense_out, scale, divisor, L, make_fastdiv((uint32_t)Ny), P.Ms};
        launch_k<KXBwd<T>>(dim3((nrows + L.R - 1) / L.R), dim3(kFftThreads), lds, st, A);
    }
}

template <class T> bool poisson_source_reads_no_halo(const PoissonPlan<T> &P, const Ctx<T> &C) {
    const DevGrid<T> &g = C.g;
    if (g.kbot || g.tx == OM_FLAT || g.ty == OM_FLAT || g.tz == OM_FLAT || (g.Nx & 1) || P.tridiagonal) return false;
    const int M = g.Nx / 2;
    const int R = fft_lines_per_block(M, M, sizeof(C2<T>));
    return R >= 1 && g.Nx <= kXFwdFastElems * (kFftThreads / R);
}

#ifndef OM_ONLY_DOUBLE
template bool poisson_source_reads_no_halo<float>(const PoissonPlan<float> &, const Ctx<float> &);
template void launch_poisson_residual<float>(const Ctx<float> &, const RefineOp &, const float *, int64_t, int64_t,
                                             const double *, double, float *, hipStream_t);
template void launch_add_correction<float>(const DevGrid<float> &, float *, int64_t, int64_t, const float *, hipStream_t);
template void launch_poisson<float>(const PoissonPlan<float> &, const Ctx<float> &, const float *, float *, int, float,
                                    hipStream_t);
#endif
#ifndef OM_ONLY_FLOAT
template bool poisson_source_reads_no_halo<double>(const PoissonPlan<double> &, const Ctx<double> &);
template void launch_poisson_residual<double>(const Ctx<double> &, const RefineOp &, const double *, int64_t, int64_t,
                                              const double *, double, double *, hipStream_t);
template void launch_add_correction<double>(const DevGrid<double> &, double *, int64_t, int64_t, const double *, hipStream_t);
template void launch_poisson<double>(const PoissonPlan<double> &, const Ctx<double> &, const double *, double *, int, double,
                                     hipStream_t);
#endif

}  // namespace om
