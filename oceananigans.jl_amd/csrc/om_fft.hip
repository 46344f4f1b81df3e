// om_fft.hip — pressure Poisson solvers on gfx950.
//
// Spectral array S (complex, x fastest): S[(k·Ny + j)·Mx + kx] with Mx = Nx/2 + 1 (real-to-complex
// along a Periodic x) or 1 (Flat x).  Passes, each one HBM read + one write of S:
//   X-forward  : source term (∇·U* formed on the fly, optionally ×Δzᶜ) → real-to-complex FFT
//   Y-forward  : complex FFT along y                    (skipped when y is the last axis)
//   last axis  : forward transform · eigenvalue division · backward transform in one kernel
//                (Bounded z: the Neumann problem on N points is the periodic problem on the
//                 2N-point even extension, so the z line is mirrored in LDS and solved there;
//                 Fourier–tridiagonal: Thomas algorithm along z)
//   Y-backward, X-backward (complex-to-real, written into the pressure field's interior)
// Each line transform is a radix-2 Stockham FFT in LDS for power-of-two lengths and a direct
// DFT otherwise.  Reference: src/Solvers/fft_based_poisson_solver.jl:94-124,
// src/Solvers/fourier_tridiagonal_poisson_solver.jl:231-260, src/Solvers/batched_tridiagonal_solver.jl:211-243.
#include <hip/hip_runtime.h>

#include "om_internal.h"
#include "om_kernels.h"

namespace om {

template <class T> using C2 = typename Vec2<T>::type;

template <class T> __device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { C2<T> r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <class T> __device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { C2<T> r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <class T> __device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) { C2<T> r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r; }
template <class T> __device__ __forceinline__ C2<T> cconj(C2<T> a) { a.y = -a.y; return a; }
template <class T> __device__ __forceinline__ C2<T> cmk(T x, T y) { C2<T> r; r.x = x; r.y = y; return r; }

// R lines of length n stored [r][n] in buffer a; b is scratch of the same size.
// tw[m] = exp(-2πi m / n), m < n.  Returns the buffer that holds the transform (unnormalised).
template <class T>
__device__ C2<T> *lds_fft(C2<T> *a, C2<T> *b, int R, int n, const C2<T> *__restrict__ tw, bool inverse) {
    const int nt = blockDim.x * blockDim.y;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    if ((n & (n - 1)) == 0) {
        // power of two: Stockham stages of radix 4 (one radix-2 stage first when log2 n is odd),
        // half as many LDS passes and barriers as radix 2.  A radix-R stage at sub-length m maps
        // a[kk + jj·m + p·n/R] (p < R) to b[kk + R·jj·m + q·m] = w^{q·jj·m} Σ_p c_p e^{∓2πi pq/R}.
        // Every index split is a shift and a mask (an integer division by a runtime value costs
        // tens of instructions per butterfly on CDNA).
        const int lg = __builtin_ctz((unsigned)n);
        int lm = 0;   // log2 of the current sub-length m
        if (lg & 1) {
            const int half = n >> 1, lh = lg - 1;
            for (int t = tid; t < R * half; t += nt) {
                const int r = t >> lh, jj = t & (half - 1);
                const int base = (r << lg) + jj;
                const C2<T> c0 = a[base];
                const C2<T> c1 = a[base + half];
                C2<T> w = tw[jj];
                if (inverse) w.y = -w.y;
                const int ob = (r << lg) + (jj << 1);
                b[ob] = cadd<T>(c0, c1);
                b[ob + 1] = cmul<T>(w, csub<T>(c0, c1));
            }
            __syncthreads();
            C2<T> *tmp = a; a = b; b = tmp;
            lm = 1;
        }
        const int quarter = n >> 2, lq = lg - 2;
        for (; lm < lg; lm += 2) {
            const int m = 1 << lm;
            for (int t = tid; t < R * quarter; t += nt) {
                const int r = t >> lq, bf = t & (quarter - 1);
                const int jj = bf >> lm, kk = bf & (m - 1);
                const int ib = (r << lg) + kk + (jj << lm);
                const C2<T> c0 = a[ib], c1 = a[ib + quarter], c2 = a[ib + 2 * quarter], c3 = a[ib + 3 * quarter];
                const C2<T> t0 = cadd<T>(c0, c2), t1 = csub<T>(c0, c2), t2 = cadd<T>(c1, c3), d = csub<T>(c1, c3);
                const C2<T> t3 = inverse ? cmk<T>(-d.y, d.x) : cmk<T>(d.y, -d.x);   // ±i·(c1 − c3)
                const int e = jj << lm;
                C2<T> w1 = tw[e], w2 = tw[2 * e], w3 = tw[3 * e];
                if (inverse) { w1.y = -w1.y; w2.y = -w2.y; w3.y = -w3.y; }
                const int ob = (r << lg) + kk + (jj << (lm + 2));
                b[ob] = cadd<T>(t0, t2);
                b[ob + m] = cmul<T>(w1, cadd<T>(t1, t3));
                b[ob + 2 * m] = cmul<T>(w2, csub<T>(t0, t2));
                b[ob + 3 * m] = cmul<T>(w3, csub<T>(t1, t3));
            }
            __syncthreads();
            C2<T> *tmp = a; a = b; b = tmp;
        }
        return a;
    }
    for (int t = tid; t < R * n; t += nt) {
        const int r = t / n, ko = t - r * n;
        C2<T> s = cmk<T>(T(0), T(0));
        for (int q = 0; q < n; ++q) {
            C2<T> w = tw[(int)(((int64_t)ko * q) % n)];
            if (inverse) w.y = -w.y;
            s = cadd<T>(s, cmul<T>(a[r * n + q], w));
        }
        b[t] = s;
    }
    __syncthreads();
    return b;
}

// ---------------------------------------------------------------------------------------------
// X-forward: source term → R2C along x.  One block handles R rows (j, k).
// src == nullptr: the source is ∇·U of the state (solve_for_pressure.jl:12-18); else a dense
// interior array.  scale_dzc: multiply by Δzᶜ(k) (Fourier–tridiagonal source, :36-42).
// ---------------------------------------------------------------------------------------------
template <class T> struct XFwdArgs {
    Ctx<T> C;
    const T *src;
    C2<T> *S;
    const C2<T> *tw;      // length Nx/2 table (for the half-length complex FFT)
    const C2<T> *post;    // exp(-2πi k / Nx), k = 0 .. Nx/2
    int scale_dzc;
    int R;
    FastDiv fNx, fNy, fMx;
};

template <class T> __device__ __forceinline__ T source_at(const Ctx<T> &C, const T *src, int i, int j, int k) {
    if (src) return src[(int64_t)i + (int64_t)C.g.Nx * ((int64_t)j + (int64_t)C.g.Ny * k)];
    // divergence with periodic index wrap in x and y (halos of U* are not needed)
    const int ip = (i + 1 == C.g.Nx) ? 0 : i + 1;
    const int jp = (j + 1 == C.g.Ny) ? 0 : j + 1;
    return C.divU(i, j, k, ip, jp);
}

template <class T> __global__ void __launch_bounds__(256) k_xfwd(XFwdArgs<T> A) {
    extern __shared__ unsigned char smem_raw[];
    const Ctx<T> &C = A.C;
    const int Nx = C.g.Nx, Ny = C.g.Ny, Nz = C.g.Nz;
    const bool half = (Nx & 1) == 0;
    const int M = half ? Nx / 2 : Nx;      // line length of the complex FFT
    const int Mx = Nx / 2 + 1;
    const int R = A.R;
    C2<T> *a = reinterpret_cast<C2<T> *>(smem_raw);
    C2<T> *b = a + R * M;
    const int row0 = blockIdx.x * R;
    const int nrows = Ny * Nz;
    const int tid = threadIdx.x;
    if (!half) for (int t = tid; t < R * M; t += blockDim.x) a[t] = cmk<T>(T(0), T(0));
    __syncthreads();
    for (int t = tid; t < R * Nx; t += blockDim.x) {
        const int r = (int)fdiv((uint32_t)t, A.fNx), i = t - r * Nx;
        const int row = row0 + r;
        T x = T(0);
        if (row < nrows) {
            const int k = (int)fdiv((uint32_t)row, A.fNy), j = row - k * Ny;
            x = source_at<T>(C, A.src, i, j, k);
            if (A.scale_dzc) x = C.g.dzc[k] * x;
        }
        T *ap = reinterpret_cast<T *>(a);
        if (half) ap[2 * (r * M + (i >> 1)) + (i & 1)] = x;
        else ap[2 * (r * M + i)] = x;
    }
    __syncthreads();
    C2<T> *Z = lds_fft<T>(a, b, R, M, A.tw, false);
    for (int t = tid; t < R * Mx; t += blockDim.x) {
        const int r = (int)fdiv((uint32_t)t, A.fMx), kk = t - r * Mx;
        const int row = row0 + r;
        if (row >= nrows) continue;
        if (!half) { A.S[(int64_t)row * Mx + kk] = Z[r * M + kk]; continue; }
        const C2<T> zk = Z[r * M + (kk == M ? 0 : kk)];
        const C2<T> zc = cconj<T>(Z[r * M + (kk == 0 ? 0 : M - kk)]);
        const C2<T> E = cmk<T>(T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y + zc.y));
        // O = (zk - zc) / (2i) = (-i/2)(zk - zc)
        const C2<T> d = csub<T>(zk, zc);
        const C2<T> O = cmk<T>(T(0.5) * d.y, T(-0.5) * d.x);
        A.S[(int64_t)row * Mx + kk] = cadd<T>(E, cmul<T>(A.post[kk], O));
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
    int R;
    FastDiv fNx, fNy, fM;
};

template <class T> __global__ void __launch_bounds__(256) k_xbwd(XBwdArgs<T> A) {
    extern __shared__ unsigned char smem_raw[];
    const int Nx = A.g.Nx, Ny = A.g.Ny, Nz = A.g.Nz;
    const bool half = (Nx & 1) == 0;
    const int M = half ? Nx / 2 : Nx;
    const int Mx = Nx / 2 + 1;
    const int R = A.R;
    C2<T> *a = reinterpret_cast<C2<T> *>(smem_raw);
    C2<T> *b = a + R * M;
    const int row0 = blockIdx.x * R;
    const int nrows = Ny * Nz;
    const int tid = threadIdx.x;
    for (int t = tid; t < R * M; t += blockDim.x) {
        const int r = (int)fdiv((uint32_t)t, A.fM), kk = t - r * M;
        const int row = row0 + r;
        C2<T> Zk = cmk<T>(T(0), T(0));
        if (row < nrows) {
            if (!half) {
                Zk = kk < Mx ? A.S[(int64_t)row * Mx + kk] : cconj<T>(A.S[(int64_t)row * Mx + (Nx - kk)]);
            } else {
                const C2<T> Xk = A.S[(int64_t)row * Mx + kk];
                const C2<T> Xc = cconj<T>(A.S[(int64_t)row * Mx + (M - kk)]);
                const C2<T> E = cmk<T>(T(0.5) * (Xk.x + Xc.x), T(0.5) * (Xk.y + Xc.y));
                const C2<T> d = csub<T>(Xk, Xc);
                // O = (Xk - conj(X_{M-k})) · e^{+2πik/N} / 2 ;  Z = E + i O
                const C2<T> O = cmul<T>(cmk<T>(T(0.5) * d.x, T(0.5) * d.y), cconj<T>(A.post[kk]));
                Zk = cmk<T>(E.x - O.y, E.y + O.x);
            }
        }
        a[r * M + kk] = Zk;
    }
    __syncthreads();
    C2<T> *z = lds_fft<T>(a, b, R, M, A.tw, true);
    for (int t = tid; t < R * Nx; t += blockDim.x) {
        const int r = (int)fdiv((uint32_t)t, A.fNx), i = t - r * Nx;
        const int row = row0 + r;
        if (row >= nrows) continue;
        const int k = (int)fdiv((uint32_t)row, A.fNy), j = row - k * Ny;
        T x;
        if (half) { const C2<T> zz = z[r * M + (i >> 1)]; x = (i & 1) ? zz.y : zz.x; }
        else x = z[r * M + i].x;
        x = (x * A.scale) / A.divisor;
        if (A.dense) A.out[(int64_t)i + (int64_t)Nx * ((int64_t)j + (int64_t)Ny * k)] = x;
        else A.out[idx3(i, j, k, A.g.sy, A.g.sz)] = x;
    }
}

// ---------------------------------------------------------------------------------------------
// Y pass: complex FFT along y for R consecutive kx columns at fixed k.
// fused_divide: forward → divide by eigenvalues → backward (y is the last non-Flat axis).
// ---------------------------------------------------------------------------------------------
template <class T> struct YArgs {
    DevGrid<T> g;
    C2<T> *S;
    const C2<T> *tw;
    int Mx;
    int inverse;
    int fused_divide;
    const T *lx, *ly;   // eigenvalues (fused)
    int R, lgR;         // rows per block (a power of two) and its log2
};

template <class T> __global__ void __launch_bounds__(256) k_ypass(YArgs<T> A) {
    extern __shared__ unsigned char smem_raw[];
    const int Ny = A.g.Ny, Mx = A.Mx, R = A.R;
    C2<T> *a = reinterpret_cast<C2<T> *>(smem_raw);
    C2<T> *b = a + R * Ny;
    const int kx0 = blockIdx.x * R;
    const int k = blockIdx.y;
    const int tid = threadIdx.x;
    for (int t = tid; t < R * Ny; t += blockDim.x) {
        const int r = t & (R - 1), j = t >> A.lgR;
        const int kx = kx0 + r;
        a[r * Ny + j] = kx < Mx ? A.S[((int64_t)k * Ny + j) * Mx + kx] : cmk<T>(T(0), T(0));
    }
    __syncthreads();
    C2<T> *res = lds_fft<T>(a, b, R, Ny, A.tw, A.inverse != 0);
    if (A.fused_divide) {
        C2<T> *other = (res == a) ? b : a;
        for (int t = tid; t < R * Ny; t += blockDim.x) {
            const int r = t & (R - 1), j = t >> A.lgR;
            const int kx = kx0 + r;
            C2<T> v = res[r * Ny + j];
            if (kx < Mx) {
                const T den = A.lx[kx] + A.ly[j];
                v = (kx == 0 && j == 0) ? cmk<T>(T(0), T(0)) : cmk<T>(-v.x / den, -v.y / den);
            }
            res[r * Ny + j] = v;
        }
        __syncthreads();
        res = lds_fft<T>(res, other, R, Ny, A.tw, true);
    }
    for (int t = tid; t < R * Ny; t += blockDim.x) {
        const int r = t & (R - 1), j = t >> A.lgR;
        const int kx = kx0 + r;
        if (kx < Mx) A.S[((int64_t)k * Ny + j) * Mx + kx] = res[r * Ny + j];
    }
}

// ---------------------------------------------------------------------------------------------
// Z pass (last axis): forward transform, eigenvalue division, backward transform.
// Bounded z: even extension of length 2Nz (the mirrored periodic problem), Periodic z: length Nz.
// ---------------------------------------------------------------------------------------------
template <class T> struct ZArgs {
    DevGrid<T> g;
    C2<T> *S;
    const C2<T> *tw;    // table for length L
    int Mx, L;          // L = 2Nz (Bounded) or Nz (Periodic)
    const T *lx, *ly, *lz;   // lz has L entries
    int R, lgR;
};

template <class T> __global__ void __launch_bounds__(256) k_zpass(ZArgs<T> A) {
    extern __shared__ unsigned char smem_raw[];
    const int Ny = A.g.Ny, Nz = A.g.Nz, Mx = A.Mx, R = A.R, L = A.L;
    const bool mirror = (L == 2 * Nz);
    C2<T> *a = reinterpret_cast<C2<T> *>(smem_raw);
    C2<T> *b = a + R * L;
    const int kx0 = blockIdx.x * R;
    const int j = blockIdx.y;
    const int tid = threadIdx.x;
    for (int t = tid; t < R * Nz; t += blockDim.x) {
        const int r = t & (R - 1), k = t >> A.lgR;
        const int kx = kx0 + r;
        const C2<T> v = kx < Mx ? A.S[((int64_t)k * Ny + j) * Mx + kx] : cmk<T>(T(0), T(0));
        a[r * L + k] = v;
        if (mirror) a[r * L + (L - 1 - k)] = v;
    }
    __syncthreads();
    C2<T> *res = lds_fft<T>(a, b, R, L, A.tw, false);
    C2<T> *other = (res == a) ? b : a;
    for (int t = tid; t < R * L; t += blockDim.x) {
        const int r = t & (R - 1), k = t >> A.lgR;
        const int kx = kx0 + r;
        C2<T> v = res[r * L + k];
        if (kx < Mx) {
            const T den = (A.lx[kx] + A.ly[j]) + A.lz[k];
            v = (kx == 0 && j == 0 && k == 0) ? cmk<T>(T(0), T(0)) : cmk<T>(-v.x / den, -v.y / den);
        }
        res[r * L + k] = v;
    }
    __syncthreads();
    res = lds_fft<T>(res, other, R, L, A.tw, true);
    for (int t = tid; t < R * Nz; t += blockDim.x) {
        const int r = t & (R - 1), k = t >> A.lgR;
        const int kx = kx0 + r;
        if (kx < Mx) A.S[((int64_t)k * Ny + j) * Mx + kx] = res[r * L + k];
    }
}

// ---------------------------------------------------------------------------------------------
// Fourier–tridiagonal z solve (Thomas per (kx, ky) column; diagonals from
// fourier_tridiagonal_poisson_solver.jl:199-229).  One thread per column; t[k] kept in LDS.
// The (0,0) column is singular: its last pivot is a rounding residue and is elided as in the
// reference; the constant this leaves is removed with the column mean (ϕ -= mean(ϕ), :253).
// ---------------------------------------------------------------------------------------------
template <class T> struct TriArgs {
    DevGrid<T> g;
    C2<T> *S;
    int Mx;
    const T *lx, *ly;
    T eps10;
    T zscale;   // (1 / L_backward_norm) applied later; unused here
};

template <class T> __global__ void __launch_bounds__(64) k_tridiag(TriArgs<T> A) {
    extern __shared__ unsigned char smem_raw[];
    T *tt = reinterpret_cast<T *>(smem_raw);   // [Nz][blockDim.x]
    const int Ny = A.g.Ny, Nz = A.g.Nz, Mx = A.Mx;
    const int bw = blockDim.x;
    const int kx = blockIdx.x * bw + threadIdx.x;
    const int j = blockIdx.y;
    if (kx >= Mx) return;
    const T lam = A.lx[kx] + A.ly[j];
    const T *idzf = A.g.idzf;
    auto diag = [&](int k) -> T {   // 0-based k
        const T dzc = A.g.dzc[k];
        if (Nz == 1) return -dzc * lam;
        if (k == 0) return -idzf[1] - dzc * lam;
        if (k == Nz - 1) return -idzf[Nz - 1] - dzc * lam;
        return -(idzf[k + 1] + idzf[k]) - dzc * lam;
    };
    auto off = [&](int k) -> T { return idzf[k + 1]; };   // a[k] = c[k] = 1/Δzᶠ(k+1) (reference q = k+1)
    auto at = [&](int k) -> int64_t { return ((int64_t)k * Ny + j) * Mx + kx; };
    T beta = diag(0);
    C2<T> f = A.S[at(0)];
    C2<T> phi = cmk<T>(f.x / beta, f.y / beta);
    A.S[at(0)] = phi;
    for (int k = 1; k < Nz; ++k) {
        const T c = off(k - 1), a = off(k - 1);
        const T t = c / beta;
        tt[k * bw + threadIdx.x] = t;
        beta = diag(k) - a * t;
        f = A.S[at(k)];
        if (fabs(beta) > A.eps10) phi = cmk<T>((f.x - a * phi.x) / beta, (f.y - a * phi.y) / beta);
        else phi = cmk<T>(T(0), T(0));
        A.S[at(k)] = phi;
    }
    for (int k = Nz - 2; k >= 0; --k) {
        const T t = tt[(k + 1) * bw + threadIdx.x];
        C2<T> p = A.S[at(k)];
        p = cmk<T>(p.x - t * phi.x, p.y - t * phi.y);
        A.S[at(k)] = p;
        phi = p;
    }
    if (kx == 0 && j == 0) {
        // remove the mean of the (0,0) column: equals removing the volume mean of ϕ
        C2<T> s = cmk<T>(T(0), T(0));
        for (int k = 0; k < Nz; ++k) s = cadd<T>(s, A.S[at(k)]);
        s = cmk<T>(s.x / T(Nz), s.y / T(Nz));
        for (int k = 0; k < Nz; ++k) A.S[at(k)] = csub<T>(A.S[at(k)], s);
    }
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
static int rows_for(int n, size_t elem, size_t budget) {
    int r = (int)(budget / (2 * (size_t)n * elem));
    if (r < 1) return 1;
    int p = 1;
    while (p * 2 <= r) p *= 2;   // a power of two, so row indices split with shifts
    return p;
}
static int ilog2(int v) { int l = 0; while ((1 << l) < v) ++l; return l; }

template <class T> void launch_poisson(const PoissonPlan<T> &P, const Ctx<T> &C, const T *src, T *out, int dense_out,
                                       T divisor, hipStream_t st) {
    const DevGrid<T> &g = C.g;
    const size_t e = sizeof(C2<T>);
    const size_t budget = 64 * 1024;
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const int M = (Nx & 1) == 0 ? Nx / 2 : Nx;
    const int Mx = Nx / 2 + 1;
    const int nrows = Ny * Nz;
    auto rows = [&](int n) { int r = rows_for(n, e, budget); return r > 16 ? 16 : r; };
    {
        XFwdArgs<T> A{C, src, P.S, P.tw_xh, P.post, P.tridiagonal, rows(M),
                      make_fastdiv((uint32_t)Nx), make_fastdiv((uint32_t)Ny), make_fastdiv((uint32_t)Mx)};
        hipLaunchKernelGGL(k_xfwd<T>, dim3((nrows + A.R - 1) / A.R), dim3(256), 2 * A.R * M * e, st, A);
    }
    const bool z_active = g.tz != OM_FLAT;
    {
        YArgs<T> A{g, P.S, P.tw_y, Mx, 0, z_active ? 0 : 1, P.lx, P.ly, rows(Ny), ilog2(rows(Ny))};
        hipLaunchKernelGGL(k_ypass<T>, dim3((Mx + A.R - 1) / A.R, Nz), dim3(256), 2 * A.R * Ny * e, st, A);
    }
    T scale = T(1) / (T(M) * T(Ny));
    if (z_active) {
        if (P.tridiagonal) {
            TriArgs<T> A{g, P.S, Mx, P.lx, P.ly, P.eps10, T(1)};
            int bw = 64;   // columns per block, limited by the LDS holding t[k] (≤ 64 KiB)
            while (bw > 1 && (size_t)Nz * bw * sizeof(T) > budget) bw >>= 1;
            hipLaunchKernelGGL(k_tridiag<T>, dim3((Mx + bw - 1) / bw, Ny), dim3(bw), (size_t)Nz * bw * sizeof(T), st, A);
        } else {
            ZArgs<T> A{g, P.S, P.tw_z, Mx, P.Lz_line, P.lx, P.ly, P.lz, rows(P.Lz_line), ilog2(rows(P.Lz_line))};
            hipLaunchKernelGGL(k_zpass<T>, dim3((Mx + A.R - 1) / A.R, Ny), dim3(256), 2 * A.R * P.Lz_line * e, st, A);
            scale = T(1) / ((T(M) * T(Ny)) * T(P.Lz_line));
        }
        YArgs<T> A{g, P.S, P.tw_y, Mx, 1, 0, P.lx, P.ly, rows(Ny), ilog2(rows(Ny))};
        hipLaunchKernelGGL(k_ypass<T>, dim3((Mx + A.R - 1) / A.R, Nz), dim3(256), 2 * A.R * Ny * e, st, A);
    }
    {
        XBwdArgs<T> A{g, P.S, P.tw_xh, P.post, out, dense_out, scale, divisor, rows(M),
                      make_fastdiv((uint32_t)Nx), make_fastdiv((uint32_t)Ny), make_fastdiv((uint32_t)M)};
        hipLaunchKernelGGL(k_xbwd<T>, dim3((nrows + A.R - 1) / A.R), dim3(256), 2 * A.R * M * e, st, A);
    }
}

#ifndef OM_ONLY_DOUBLE
template void launch_poisson<float>(const PoissonPlan<float> &, const Ctx<float> &, const float *, float *, int, float,
                                    hipStream_t);
#endif
#ifndef OM_ONLY_FLOAT
template void launch_poisson<double>(const PoissonPlan<double> &, const Ctx<double> &, const double *, double *, int, double,
                                     hipStream_t);
#endif

}  // namespace om
