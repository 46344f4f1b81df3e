// om_learned.hip — the learned convolutional correction to the FFT preconditioner of the
// conjugate-gradient Poisson solver, and the kernels that train it.  The reference has no such
// network; it plugs in where the reference applies its FFT preconditioner
// (conjugate_gradient_poisson_solver.jl:169-211) and is used inside flexible CG.
//
// Network: a four-level hierarchy of seven 3×3×3 convolutions, 16 → 16 channels, x and y periodic,
// z zero-padded at every level:
//     a0 = relu(conv0(x))            level 0 (the grid)      a3 = relu(conv3(pool(a2)))   level 3
//     a1 = relu(conv1(pool(a0)))     level 1 (2× coarser)    b2 = relu(conv4(a2 + up(a3)))
//     a2 = relu(conv2(pool(a1)))     level 2                 b1 = relu(conv5(a1 + up(b2)))
//     out = conv6(a0 + up(b1)), channel 0
// pool is the mean of the 2×2×2 children, up copies a coarse value to its children.  The coarse
// levels reach 8× farther than the grid's 3×3×3 stencils: the large-scale part of the FFT
// preconditioner's error near the immersed bottom is what sets the CG iteration count.  The first
// layer's input carries 3 meaningful channels.  Weights, activations and back-propagated
// gradients are BF16; products accumulate in FP32 on the matrix cores (v_mfma_f32_16x16x32_bf16),
// pooled and summed values in FP32 before rounding.
//
// Layout.  Activations are channel-last BF16 in a padded box with one halo cell on each side:
// voxel (i, j, k), i ∈ [−1, nx], … at ((k + 1)·(ny + 2) + j + 1)·(nx + 2) + i + 1, 16 channels
// (32 bytes) per voxel.  A wave computes 16 consecutive voxels of one x row × 16 output channels
// as D = A·B with A = weights (16 cout × 32 k) and B = activations (32 k × 16 voxels); the K = 27·16
// contraction runs over 14 MFMA steps of two taps × 16 input channels (the 28th tap is zero).
// Lane l supplies k = 8·(l / 16) + e, e < 8, i.e. tap 2s + l/32 and input channels
// 8·((l/16) mod 2) + e — one 16-byte load per step — and receives output channels
// 4·(l / 16) + e, e < 4, of voxel l mod 16.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "om_learned.h"

#ifndef OM_ONLY_DOUBLE   // precision-independent: compiled into the FP32 object only
namespace om {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int64_t pidx(const NetDims &d, int i, int j, int k) {
    return ((int64_t)(k + 1) * (d.ny + 2) + (j + 1)) * (d.nx + 2) + (i + 1);
}
__device__ __forceinline__ void tap_offsets(int t, int &dx, int &dy, int &dz) {
    dz = t / 9 - 1;
    dy = (t / 3) % 3 - 1;
    dx = t % 3 - 1;
}

// -------------------------------------------------------------------------------------------
// Convolution on the matrix cores.  One wave = 16 voxels × 16 output channels.
// EPI_RELU: y = bf16(max(0, acc)) into a padded activation;  EPI_OUT: channel 0 as FP32 dense;
// EPI_MASK: y = bf16(acc · [act > 0]) (the back-propagated gradient through the ReLU of `act`);
// EPI_IN: channels 0..2 as FP32 dense (the input gradient); EPI_LIN: y = bf16(acc).
// -------------------------------------------------------------------------------------------
enum { EPI_RELU = 0, EPI_OUT = 1, EPI_MASK = 2, EPI_IN = 3, EPI_LIN = 4 };

template <int EPI>
__global__ void __launch_bounds__(256) k_conv(NetDims d, const bf16 *__restrict__ x, const bf16x8 *__restrict__ wp,
                                              bf16 *__restrict__ ybf, float *__restrict__ yf, const bf16 *__restrict__ act) {
    const int lane = threadIdx.x & 63;
    const int seg = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int spr = d.nx >> 4;
    const int row = seg / spr;
    const int x0 = (seg - row * spr) << 4;
    const int j = row % d.ny, k = row / d.ny;
    if (k >= d.nz) return;   // whole wave
    const int vi = x0 + (lane & 15);
    const int half = lane >> 5;                 // which of the step's two taps
    const int c0 = ((lane >> 4) & 1) << 3;      // first input channel of this lane
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int s = 0; s < 14; ++s) {
        const int t = 2 * s + half;
        bf16x8 b;
        if (t < 27) {
            int dx, dy, dz;
            tap_offsets(t, dx, dy, dz);
            b = *reinterpret_cast<const bf16x8 *>(x + pidx(d, vi + dx, j + dy, k + dz) * 16 + c0);
        } else {
            for (int e = 0; e < 8; ++e) b[e] = (bf16)0.0f;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wp[s * 64 + lane], b, acc, 0, 0, 0);
    }
    const int cout0 = (lane >> 4) << 2;
    const int64_t o = pidx(d, vi, j, k);
    if (EPI == EPI_RELU || EPI == EPI_MASK || EPI == EPI_LIN) {
        bf16x4 out;
        if (EPI == EPI_MASK) {
            const bf16x4 a = *reinterpret_cast<const bf16x4 *>(act + o * 16 + cout0);
            for (int e = 0; e < 4; ++e) out[e] = (bf16)((float)a[e] > 0.f ? acc[e] : 0.f);
        } else if (EPI == EPI_LIN) {
            for (int e = 0; e < 4; ++e) out[e] = (bf16)acc[e];
        } else {
            for (int e = 0; e < 4; ++e) out[e] = (bf16)(acc[e] > 0.f ? acc[e] : 0.f);
        }
        *reinterpret_cast<bf16x4 *>(ybf + o * 16 + cout0) = out;
    } else {
        const int64_t dn = (int64_t)vi + (int64_t)d.nx * ((int64_t)j + (int64_t)d.ny * k);
        if (cout0 == 0) {
            if (EPI == EPI_OUT) yf[dn] = acc[0];
            else { yf[dn * 3 + 0] = acc[0]; yf[dn * 3 + 1] = acc[1]; yf[dn * 3 + 2] = acc[2]; }
        }
    }
}

// -------------------------------------------------------------------------------------------
// The same convolution with the input staged in LDS.  A block of 8 waves computes a tile of
// 16 (x) × 8 (y) × 8 (z) voxels: its 18 × 10 × 10 input voxels (one-cell halo, 16 channels,
// 57.6 KB) are loaded from the padded activation once, instead of once per tap and wave (1.76×
// the voxels the tile produces; a 16 × 4 × 4 tile stages 2.5×), and each wave (one z level of the
// tile) reuses every weight fragment for its eight 16-voxel rows.  Same operands, same MFMA
// sequence per output: the results equal k_conv's bit for bit.
// -------------------------------------------------------------------------------------------
constexpr int CT_X = 16, CT_Y = 4, CT_Z = 4;   // the weight-gradient kernel's tile (below)
constexpr int CT_SX = CT_X + 2, CT_SY = CT_Y + 2, CT_SZ = CT_Z + 2;
constexpr int CT_V = CT_SX * CT_SY * CT_SZ;   // staged voxels per tile
constexpr int CV_X = 16, CV_Y = 8, CV_Z = 8, CV_THREADS = 64 * CV_Z;   // the convolution's tile
constexpr int CV_SX = CV_X + 2, CV_SY = CV_Y + 2, CV_SZ = CV_Z + 2;
constexpr int CV_V = CV_SX * CV_SY * CV_SZ;

template <int EPI>
__global__ void __launch_bounds__(CV_THREADS) k_conv_tiled(NetDims d, const bf16 *__restrict__ x, const bf16x8 *__restrict__ wp,
                                                           bf16 *__restrict__ ybf, float *__restrict__ yf, const bf16 *__restrict__ act) {
    __shared__ bf16x8 tile[CV_V * 2];
    const int ntx = d.nx / CV_X, nty = (d.ny + CV_Y - 1) / CV_Y;
    const int b = blockIdx.x;
    const int tx = b % ntx, rest = b / ntx;
    const int ty = rest % nty, tz = rest / nty;
    const int i0 = tx * CV_X, j0 = ty * CV_Y, k0 = tz * CV_Z;
    // stage the tile: voxel (xx, yy, zz) of the tile is padded voxel (i0 − 1 + xx, j0 − 1 + yy, k0 − 1 + zz);
    // rows and levels past the padded box's last (ny, nz) are staged as zeros and never used
    for (int q = threadIdx.x; q < CV_V * 2; q += CV_THREADS) {
        const int v = q >> 1, h = q & 1;
        const int xx = v % CV_SX, r = v / CV_SX;
        const int yy = r % CV_SY, zz = r / CV_SY;
        const int i = i0 - 1 + xx, j = j0 - 1 + yy, k = k0 - 1 + zz;
        bf16x8 val;
        if (j <= d.ny && k <= d.nz) {
            val = *reinterpret_cast<const bf16x8 *>(x + pidx(d, i, j, k) * 16 + h * 8);
        } else {
            for (int e = 0; e < 8; ++e) val[e] = (bf16)0.0f;
        }
        tile[q] = val;
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = k0 + w;
    if (k >= d.nz) return;   // whole wave
    const int vi = lane & 15;
    const int half = lane >> 5;                 // which of the step's two taps
    const int c0h = (lane >> 4) & 1;            // which 8 input channels
    f32x4 acc[CV_Y];
#pragma unroll
    for (int sgm = 0; sgm < CV_Y; ++sgm) acc[sgm] = f32x4{0.f, 0.f, 0.f, 0.f};
    bf16x8 zero;
    for (int e = 0; e < 8; ++e) zero[e] = (bf16)0.0f;
#pragma unroll 2
    for (int s = 0; s < 14; ++s) {
        const int t = 2 * s + half;
        int dx = 0, dy = 0, dz = 0;
        if (t < 27) tap_offsets(t, dx, dy, dz);
        const bf16x8 A = wp[s * 64 + lane];
        const int base = ((w + 1 + dz) * CV_SY + (1 + dy)) * CV_SX + (vi + 1 + dx);
#pragma unroll
        for (int sgm = 0; sgm < CV_Y; ++sgm) {
            const bf16x8 Bv = t < 27 ? tile[(base + sgm * CV_SX) * 2 + c0h] : zero;
            acc[sgm] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, Bv, acc[sgm], 0, 0, 0);
        }
    }
    const int cout0 = (lane >> 4) << 2;
#pragma unroll
    for (int sgm = 0; sgm < CV_Y; ++sgm) {
        const int j = j0 + sgm;
        if (j >= d.ny) break;
        const int iv = i0 + vi;
        const int64_t o = pidx(d, iv, j, k);
        if (EPI == EPI_RELU || EPI == EPI_MASK || EPI == EPI_LIN) {
            bf16x4 out;
            if (EPI == EPI_MASK) {
                const bf16x4 a = *reinterpret_cast<const bf16x4 *>(act + o * 16 + cout0);
                for (int e = 0; e < 4; ++e) out[e] = (bf16)((float)a[e] > 0.f ? acc[sgm][e] : 0.f);
            } else if (EPI == EPI_LIN) {
                for (int e = 0; e < 4; ++e) out[e] = (bf16)acc[sgm][e];
            } else {
                for (int e = 0; e < 4; ++e) out[e] = (bf16)(acc[sgm][e] > 0.f ? acc[sgm][e] : 0.f);
            }
            *reinterpret_cast<bf16x4 *>(ybf + o * 16 + cout0) = out;
        } else {
            const int64_t dn = (int64_t)iv + (int64_t)d.nx * ((int64_t)j + (int64_t)d.ny * k);
            if (cout0 == 0) {
                if (EPI == EPI_OUT) yf[dn] = acc[sgm][0];
                else { yf[dn * 3 + 0] = acc[sgm][0]; yf[dn * 3 + 1] = acc[sgm][1]; yf[dn * 3 + 2] = acc[sgm][2]; }
            }
        }
    }
}

// -------------------------------------------------------------------------------------------
// Level transfers, one thread per (voxel, 8 channels); sums in FP32, in the order (z, y, x) of the
// child offsets, rounded to BF16 once.  dc: the coarse level, df: the fine level (2× each axis).
// -------------------------------------------------------------------------------------------
__device__ __forceinline__ bool level_voxel(const NetDims &d, int64_t t, int &i, int &j, int &k, int &h) {
    const int64_t nv = (int64_t)d.nx * d.ny * d.nz;
    if (t >= 2 * nv) return false;
    const int64_t v = t >> 1;
    h = (int)(t & 1);
    i = (int)(v % d.nx);
    const int64_t r = v / d.nx;
    j = (int)(r % d.ny);
    k = (int)(r / d.ny);
    return true;
}
__device__ __forceinline__ void child_sum8(const NetDims &df, const bf16 *f, int i, int j, int k, int h, float (&s)[8]) {
    for (int e = 0; e < 8; ++e) s[e] = 0.f;
    for (int cz = 0; cz < 2; ++cz)
        for (int cy = 0; cy < 2; ++cy)
            for (int cx = 0; cx < 2; ++cx) {
                const bf16x8 c = *reinterpret_cast<const bf16x8 *>(f + pidx(df, 2 * i + cx, 2 * j + cy, 2 * k + cz) * 16 + h * 8);
                for (int e = 0; e < 8; ++e) s[e] += (float)c[e];
            }
}
// coarse = pool(fine): the mean of the children
__global__ void __launch_bounds__(256) k_pool(NetDims dc, const bf16 *fine, bf16 *coarse) {
    int i, j, k, h;
    if (!level_voxel(dc, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, i, j, k, h)) return;
    const NetDims df{dc.nx * 2, dc.ny * 2, dc.nz * 2};
    float s[8];
    child_sum8(df, fine, i, j, k, h, s);
    bf16x8 o;
    for (int e = 0; e < 8; ++e) o[e] = (bf16)(s[e] * 0.125f);
    *reinterpret_cast<bf16x8 *>(coarse + pidx(dc, i, j, k) * 16 + h * 8) = o;
}
// out = skip + up(coarse) on the fine level
__global__ void __launch_bounds__(256) k_up_add(NetDims df, const bf16 *skip, const bf16 *coarse, bf16 *out) {
    int i, j, k, h;
    if (!level_voxel(df, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, i, j, k, h)) return;
    const NetDims dc{df.nx / 2, df.ny / 2, df.nz / 2};
    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(skip + pidx(df, i, j, k) * 16 + h * 8);
    const bf16x8 c = *reinterpret_cast<const bf16x8 *>(coarse + pidx(dc, i >> 1, j >> 1, k >> 1) * 16 + h * 8);
    bf16x8 o;
    for (int e = 0; e < 8; ++e) o[e] = (bf16)((float)a[e] + (float)c[e]);
    *reinterpret_cast<bf16x8 *>(out + pidx(df, i, j, k) * 16 + h * 8) = o;
}
// the adjoint of up, through the ReLU of the coarse activation: out = Σ children(gfine) · [act > 0]
__global__ void __launch_bounds__(256) k_down_mask(NetDims dc, const bf16 *gfine, const bf16 *act, bf16 *out) {
    int i, j, k, h;
    if (!level_voxel(dc, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, i, j, k, h)) return;
    const NetDims df{dc.nx * 2, dc.ny * 2, dc.nz * 2};
    float s[8];
    child_sum8(df, gfine, i, j, k, h, s);
    const int64_t o = pidx(dc, i, j, k) * 16 + h * 8;
    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(act + o);
    bf16x8 r;
    for (int e = 0; e < 8; ++e) r[e] = (bf16)((float)a[e] > 0.f ? s[e] : 0.f);
    *reinterpret_cast<bf16x8 *>(out + o) = r;
}
// a fine activation's gradient from its two uses (the skip sum and the pooled input of the next
// level, the adjoint of the mean), through its ReLU: out = (gskip + up(gcoarse)/8) · [act > 0]
__global__ void __launch_bounds__(256) k_up_combine(NetDims df, const bf16 *gskip, const bf16 *gcoarse, const bf16 *act, bf16 *out) {
    int i, j, k, h;
    if (!level_voxel(df, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, i, j, k, h)) return;
    const NetDims dc{df.nx / 2, df.ny / 2, df.nz / 2};
    const int64_t o = pidx(df, i, j, k) * 16 + h * 8;
    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(gskip + o);
    const bf16x8 c = *reinterpret_cast<const bf16x8 *>(gcoarse + pidx(dc, i >> 1, j >> 1, k >> 1) * 16 + h * 8);
    const bf16x8 m = *reinterpret_cast<const bf16x8 *>(act + o);
    bf16x8 r;
    for (int e = 0; e < 8; ++e) r[e] = (bf16)((float)m[e] > 0.f ? (float)a[e] + 0.125f * (float)c[e] : 0.f);
    *reinterpret_cast<bf16x8 *>(out + o) = r;
}

// halo of a padded activation: x, y periodic (copied from the opposite interior edge), z zero.
// Threads enumerate the halo voxels only, two per voxel (8 channels each): the two z faces (all
// of their (nx + 2)(ny + 2) voxels, zeros), then the two y faces of the interior levels, then the
// two x faces of the interior levels and rows.
__global__ void __launch_bounds__(256) k_net_halo(NetDims d, bf16 *a) {
    const int64_t px = d.nx + 2, py = d.ny + 2;
    const int64_t nzf = 2 * px * py, nyf = 2 * (int64_t)d.nz * px, nxf = 2 * (int64_t)d.nz * d.ny;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t v = t >> 1;
    const int hh = (int)(t & 1);
    int i, j, k;
    bool zero = false;
    if (v < nzf) {
        const int64_t per = px * py;
        const bool top = v >= per;
        const int64_t r = top ? v - per : v;
        k = top ? d.nz : -1;
        j = (int)(r / px) - 1;
        i = (int)(r - (r / px) * px) - 1;
        zero = true;
    } else if (v < nzf + nyf) {
        const int64_t per = (int64_t)d.nz * px;
        int64_t u = v - nzf;
        const bool hi = u >= per;
        if (hi) u -= per;
        j = hi ? d.ny : -1;
        k = (int)(u / px);
        i = (int)(u - (int64_t)k * px) - 1;
    } else if (v < nzf + nyf + nxf) {
        const int64_t per = (int64_t)d.nz * d.ny;
        int64_t u = v - nzf - nyf;
        const bool hi = u >= per;
        if (hi) u -= per;
        i = hi ? d.nx : -1;
        k = (int)(u / d.ny);
        j = (int)(u - (int64_t)k * d.ny);
    } else {
        return;
    }
    bf16x8 val;
    if (zero) {
        for (int e = 0; e < 8; ++e) val[e] = (bf16)0.0f;
    } else {
        const int ii = (i + d.nx) % d.nx, jj = (j + d.ny) % d.ny;
        val = *reinterpret_cast<const bf16x8 *>(a + pidx(d, ii, jj, k) * 16 + hh * 8);
    }
    *reinterpret_cast<bf16x8 *>(a + pidx(d, i, j, k) * 16 + hh * 8) = val;
}

// weight fragments for the MFMA A operand: wp[s][lane][e] = W[t][c][o] with t = 2s + lane/32,
// c = 8·((lane/16) mod 2) + e, o = lane mod 16 (forward), or the adjoint's W[26 − t][o][c]
// (flipped taps, transposed channels); FP32 master weights W[layer][t][c][o]
__global__ void __launch_bounds__(256) k_pack_weights(const float *w, bf16 *wf, bf16 *wb) {
    const int n = NET_LAYERS * 14 * 64 * 8;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int e = q & 7, lane = (q >> 3) & 63, s = (q >> 9) % 14, l = q / (14 * 512);
    const int t = 2 * s + (lane >> 5), c = (((lane >> 4) & 1) << 3) + e, o = lane & 15;
    const float *W = w + (size_t)l * NET_WPL;
    float f = 0.f, b = 0.f;
    if (t < 27) {
        f = W[(t * 16 + c) * 16 + o];
        b = W[((26 - t) * 16 + o) * 16 + c];
    }
    wf[q] = (bf16)f;
    wb[q] = (bf16)b;
}

// -------------------------------------------------------------------------------------------
// Weight gradient dW[t][c][o] = Σ_v X[v + tap_t][c] · G[v][o] in FP32: block (tap, chunk) stages
// 64 voxels of X (shifted by the tap) and G in LDS; thread (c, o) accumulates; the chunks'
// partial sums are added in a fixed order by k_wgrad_sum.
// -------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_wgrad(NetDims d, const bf16 *x, const bf16 *g, int nchunk, float *partial) {
    __shared__ float sx[64][17], sg[64][17];
    const int t = blockIdx.x % 27, chunk = blockIdx.x / 27;
    int dx, dy, dz;
    tap_offsets(t, dx, dy, dz);
    const int64_t nv = (int64_t)d.nx * d.ny * d.nz;
    const int64_t per = (nv + nchunk - 1) / nchunk;
    const int64_t v0 = chunk * per, v1 = v0 + per < nv ? v0 + per : nv;
    const int c = threadIdx.x >> 4, o = threadIdx.x & 15;
    float acc = 0.f;
    for (int64_t vb = v0; vb < v1; vb += 64) {
        __syncthreads();
        for (int q = threadIdx.x; q < 64 * 16; q += 256) {
            const int vv = q >> 4, ch = q & 15;
            const int64_t v = vb + vv;
            float a = 0.f, b = 0.f;
            if (v < v1) {
                const int i = (int)(v % d.nx);
                const int64_t r = v / d.nx;
                const int j = (int)(r % d.ny), k = (int)(r / d.ny);
                a = (float)x[pidx(d, i + dx, j + dy, k + dz) * 16 + ch];
                b = (float)g[pidx(d, i, j, k) * 16 + ch];
            }
            sx[vv][ch] = a;
            sg[vv][ch] = b;
        }
        __syncthreads();
#pragma unroll 8
        for (int vv = 0; vv < 64; ++vv) acc += sx[vv][c] * sg[vv][o];
    }
    partial[((size_t)chunk * 27 + t) * 256 + threadIdx.x] = acc;
}
// -------------------------------------------------------------------------------------------
// The same weight gradient on the matrix cores.  For tap t, dW_t (16 c × 16 o) = Σ_v X[v + tap_t]ᵀ G[v]
// is a product with the voxels as the contraction: D = A·B with A[c][k] = X[v_k + tap][c] and
// B[k][o] = G[v_k][o], 32 voxels per MFMA (two 16-voxel x rows).  A block of 4 waves walks tiles of
// 16 × 4 × 4 voxels (grid-stride, a fixed order per block): it stages the tile's 18 × 6 × 6 input
// voxels and its 256 gradient voxels in LDS transposed to channel-major rows, so that a lane's 8
// contraction entries (8 consecutive x of one channel) are adjacent; wave w accumulates the taps
// t ≡ w (mod 4) in registers over all its tiles and writes its 16×16 partial sums per tap.  The
// block partials are summed in a fixed order by k_wgrad_sum.
// -------------------------------------------------------------------------------------------
constexpr int WG_BLOCKS_MAX = 512;
constexpr int WG_XW = 18;   // halo-tile row width

__global__ void __launch_bounds__(256) k_wgrad_mfma(NetDims d, const bf16 *__restrict__ x, const bf16 *__restrict__ g, float *partial) {
    __shared__ bf16 Xt[16][CT_SZ][CT_SY][WG_XW];   // input halo tile, channel-major
    __shared__ bf16x8 Gt[16][CT_Z][CT_Y][2];       // gradient tile, channel-major (16 x per row)
    const int ntx = d.nx / CT_X, nty = (d.ny + CT_Y - 1) / CT_Y, ntz = (d.nz + CT_Z - 1) / CT_Z;
    const int ntiles = ntx * nty * ntz;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = lane & 15;                 // A row (input channel c) / B column (output channel o)
    const int kq = lane >> 4;                // which 8 of the 32 contraction voxels
    const int xr = (kq & 1) << 3;            // their first x in the row
    const int rsel = kq >> 1;                // which of the group's two rows
    f32x4 acc[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int tx = tile % ntx, rest = tile / ntx;
        const int ty = rest % nty, tz = rest / nty;
        const int i0 = tx * CT_X, j0 = ty * CT_Y, k0 = tz * CT_Z;
        __syncthreads();   // the previous tile's readers are done
        for (int q = threadIdx.x; q < CT_V * 2; q += 256) {
            const int v = q >> 1, h = q & 1;
            const int xx = v % CT_SX, r = v / CT_SX;
            const int yy = r % CT_SY, zz = r / CT_SY;
            const int i = i0 - 1 + xx, j = j0 - 1 + yy, k = k0 - 1 + zz;
            bf16x8 val;
            if (j <= d.ny && k <= d.nz) val = *reinterpret_cast<const bf16x8 *>(x + pidx(d, i, j, k) * 16 + h * 8);
            else for (int e = 0; e < 8; ++e) val[e] = (bf16)0.0f;
            for (int e = 0; e < 8; ++e) Xt[h * 8 + e][zz][yy][xx] = val[e];
        }
        for (int q = threadIdx.x; q < CT_X * CT_Y * CT_Z * 2; q += 256) {
            const int v = q >> 1, h = q & 1;
            const int xx = v % CT_X, r = v / CT_X;
            const int yy = r % CT_Y, zz = r / CT_Y;
            const int i = i0 + xx, j = j0 + yy, k = k0 + zz;
            bf16x8 val;
            if (j < d.ny && k < d.nz) val = *reinterpret_cast<const bf16x8 *>(g + pidx(d, i, j, k) * 16 + h * 8);
            else for (int e = 0; e < 8; ++e) val[e] = (bf16)0.0f;   // outside the grid: no contribution
            bf16 *row = reinterpret_cast<bf16 *>(&Gt[0][zz][yy][0]);
            for (int e = 0; e < 8; ++e) row[(size_t)(h * 8 + e) * (CT_Z * CT_Y * 16) + xx] = val[e];
        }
        __syncthreads();
        // the B operands of the tile's 8 groups (rows 2p, 2p + 1 at level z), reused by every tap
        bf16x8 Bv[8];
#pragma unroll
        for (int grp = 0; grp < 8; ++grp) {
            const int z = grp >> 1, y = ((grp & 1) << 1) + rsel;
            Bv[grp] = Gt[m][z][y][kq & 1];
        }
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            const int t = w + 4 * q;
            if (t >= 27) break;
            int dx, dy, dz;
            tap_offsets(t, dx, dy, dz);
#pragma unroll
            for (int grp = 0; grp < 8; ++grp) {
                const int z = grp >> 1, y = ((grp & 1) << 1) + rsel;
                const bf16 *src = &Xt[m][z + 1 + dz][y + 1 + dy][xr + 1 + dx];
                bf16x8 Av;
#pragma unroll
                for (int e = 0; e < 8; ++e) Av[e] = src[e];
                acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Av, Bv[grp], acc[q], 0, 0, 0);
            }
        }
    }
    // lane holds D[c = 4·(lane / 16) + e][o = lane mod 16]
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const int t = w + 4 * q;
        if (t >= 27) break;
        float *P = partial + ((size_t)blockIdx.x * 27 + t) * 256;
        for (int e = 0; e < 4; ++e) P[(4 * kq + e) * 16 + m] = acc[q][e];
    }
}

__global__ void __launch_bounds__(256) k_wgrad_sum(const float *partial, int nchunk, float *dw) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;   // t·256 + c·16 + o
    if (q >= 27 * 256) return;
    float s = 0.f;
    for (int ch = 0; ch < nchunk; ++ch) s += partial[(size_t)ch * 27 * 256 + q];
    dw[q] = s;
}

// -------------------------------------------------------------------------------------------
// Loss Σ_active (y − t)² / N and its gradient 2 (y − t) m / N (as the first back-propagated
// gradient, BF16 channel 0 of a padded tensor whose other channels are zero); partial sums in
// FP64, reduced in a fixed order.
// -------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_loss(NetDims d, const float *y, const float *target, const float *mask, float inv_n,
                                              bf16 *g, double *partial) {
    __shared__ double red[256];
    const int64_t nv = (int64_t)d.nx * d.ny * d.nz;
    double acc = 0;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x) {
        const float df = (y[v] - target[v]) * mask[v];
        acc += (double)df * (double)df;
        const int i = (int)(v % d.nx);
        const int64_t r = v / d.nx;
        const int j = (int)(r % d.ny), k = (int)(r / d.ny);
        bf16x8 z0, z1;
        for (int e = 0; e < 8; ++e) { z0[e] = (bf16)0.0f; z1[e] = (bf16)0.0f; }
        z0[0] = (bf16)(2.f * inv_n * df);
        bf16 *p = g + pidx(d, i, j, k) * 16;
        *reinterpret_cast<bf16x8 *>(p) = z0;
        *reinterpret_cast<bf16x8 *>(p + 8) = z1;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// the back-propagated gradient at the network output, from a dense FP32 array (channel 0)
__global__ void __launch_bounds__(256) k_out_grad(NetDims d, const float *gd, bf16 *g) {
    const int64_t nv = (int64_t)d.nx * d.ny * d.nz;
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int i = (int)(v % d.nx);
    const int64_t r = v / d.nx;
    const int j = (int)(r % d.ny), k = (int)(r / d.ny);
    bf16x8 a, b;
    for (int e = 0; e < 8; ++e) { a[e] = (bf16)0.0f; b[e] = (bf16)0.0f; }
    a[0] = (bf16)gd[v];
    bf16 *p = g + pidx(d, i, j, k) * 16;
    *reinterpret_cast<bf16x8 *>(p) = a;
    *reinterpret_cast<bf16x8 *>(p + 8) = b;
}

// Σ partial[b] in a fixed order (one block)
__global__ void __launch_bounds__(256) k_loss_sum(const double *partial, double *out) {
    __shared__ double red[256];
    double acc = 0;
    for (int b = threadIdx.x; b < NET_RED_BLOCKS; b += 256) acc += partial[b];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// Adam (Kingma & Ba 2015, Algorithm 1) on the FP32 master weights; the padded input channels of
// the first layer and the unused output channels of the last stay zero.  A non-finite gradient
// (an overflowed activation) leaves its weight and moments as they are.
__global__ void __launch_bounds__(256) k_adam(float *w, float *m, float *v, const float *g, int step, float lr, float b1, float b2,
                                              float eps) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NET_LAYERS * NET_WPL) return;
    const int l = q / NET_WPL, r = q % NET_WPL, c = (r / 16) % 16, o = r % 16;
    if ((l == 0 && c >= 3) || (l == NET_LAYERS - 1 && o >= 1)) return;
    const float gq = g[q];
    if (!isfinite(gq)) return;
    const float mq = b1 * m[q] + (1.f - b1) * gq;
    const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    const float mh = mq / (1.f - powf(b1, (float)step)), vh = vq / (1.f - powf(b2, (float)step));
    w[q] = w[q] - lr * mh / (sqrtf(vh) + eps);
}

// network input from dense FP32 channels (pairs, tests): x0[v] = (in[3v], in[3v+1], in[3v+2], 0, …)
__global__ void __launch_bounds__(256) k_net_input(NetDims d, const float *in3, bf16 *x0) {
    const int64_t nv = (int64_t)d.nx * d.ny * d.nz;
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int i = (int)(v % d.nx);
    const int64_t r = v / d.nx;
    const int j = (int)(r % d.ny), k = (int)(r / d.ny);
    bf16x8 a, b;
    for (int e = 0; e < 8; ++e) { a[e] = (bf16)0.0f; b[e] = (bf16)0.0f; }
    a[0] = (bf16)in3[3 * v];
    a[1] = (bf16)in3[3 * v + 1];
    a[2] = (bf16)in3[3 * v + 2];
    bf16 *p = x0 + pidx(d, i, j, k) * 16;
    *reinterpret_cast<bf16x8 *>(p) = a;
    *reinterpret_cast<bf16x8 *>(p + 8) = b;
}

// =============================================================================================
// host side
// =============================================================================================
static inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void net_pack(NetState &S, hipStream_t st) {
    hipLaunchKernelGGL(k_pack_weights, dim3(blocks_for(NET_LAYERS * 14 * 64 * 8, 256)), dim3(256), 0, st, (const float *)S.w,
                       (bf16 *)S.wf, (bf16 *)S.wb);
}

static void conv_launch(const NetState &S, const NetDims &d, int epi, const void *x, const void *wp, void *ybf, float *yf,
                        const void *act, hipStream_t st) {
    const bf16 *X = (const bf16 *)x;
    const bf16x8 *W = (const bf16x8 *)wp;
    bf16 *Y = (bf16 *)ybf;
    const bf16 *Ac = (const bf16 *)act;
    if (S.tiled_conv) {
        const int64_t tiles = (int64_t)(d.nx / CV_X) * ((d.ny + CV_Y - 1) / CV_Y) * ((d.nz + CV_Z - 1) / CV_Z);
        const dim3 g((unsigned)tiles), b(CV_THREADS);
        switch (epi) {
            case EPI_RELU: hipLaunchKernelGGL(k_conv_tiled<EPI_RELU>, g, b, 0, st, d, X, W, Y, yf, Ac); break;
            case EPI_OUT: hipLaunchKernelGGL(k_conv_tiled<EPI_OUT>, g, b, 0, st, d, X, W, Y, yf, Ac); break;
            case EPI_MASK: hipLaunchKernelGGL(k_conv_tiled<EPI_MASK>, g, b, 0, st, d, X, W, Y, yf, Ac); break;
            case EPI_LIN: hipLaunchKernelGGL(k_conv_tiled<EPI_LIN>, g, b, 0, st, d, X, W, Y, yf, Ac); break;
            default: hipLaunchKernelGGL(k_conv_tiled<EPI_IN>, g, b, 0, st, d, X, W, Y, yf, Ac); break;
        }
        return;
    }
    const int64_t segs = (int64_t)(d.nx / 16) * d.ny * d.nz;
    const dim3 g(blocks_for(segs, 4)), b(256);
    switch (epi) {
        case EPI_RELU: hipLaunchKernelGGL(k_conv<EPI_RELU>, g, b, 0, st, d, X, W, Y, yf, Ac); break;
        case EPI_OUT: hipLaunchKernelGGL(k_conv<EPI_OUT>, g, b, 0, st, d, X, W, Y, yf, Ac); break;
        case EPI_MASK: hipLaunchKernelGGL(k_conv<EPI_MASK>, g, b, 0, st, d, X, W, Y, yf, Ac); break;
        case EPI_LIN: hipLaunchKernelGGL(k_conv<EPI_LIN>, g, b, 0, st, d, X, W, Y, yf, Ac); break;
        default: hipLaunchKernelGGL(k_conv<EPI_IN>, g, b, 0, st, d, X, W, Y, yf, Ac); break;
    }
}

static NetDims level(const NetDims &d, int l) { return NetDims{d.nx >> l, d.ny >> l, d.nz >> l}; }

static void halo(const NetDims &d, void *a, hipStream_t st) {
    const int64_t px = d.nx + 2, py = d.ny + 2;
    const int64_t n = 2 * (2 * px * py + 2 * (int64_t)d.nz * px + 2 * (int64_t)d.nz * d.ny);   // 2 threads per halo voxel
    hipLaunchKernelGGL(k_net_halo, dim3(blocks_for(n, 256)), dim3(256), 0, st, d, (bf16 *)a);
}

// one thread per (voxel, 8 channels) of level dims d
static dim3 transfer_grid(const NetDims &d) { return dim3(blocks_for(2 * (int64_t)d.nx * d.ny * d.nz, 256)); }

size_t net_arena_bytes(const NetDims &d) {
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    // level 0: x0 a0 s0 gA gB; levels 1, 2: p a s b gA gB gC; level 3: p a gA gB
    return 5 * al(net_level_bytes(d, 0)) + 7 * al(net_level_bytes(d, 1)) + 7 * al(net_level_bytes(d, 2)) +
           4 * al(net_level_bytes(d, 3));
}

void net_place_buffers(NetState &S, void *arena) {
    char *p = static_cast<char *>(arena);
    auto take = [&](int l) {
        void *r = p;
        p += (net_level_bytes(S.d, l) + 255) & ~(size_t)255;
        return r;
    };
    NetBuffers &B = S.b;
    B.x0 = take(0); B.a0 = take(0); B.s0 = take(0); B.gA[0] = take(0); B.gB[0] = take(0); B.gC[0] = nullptr;
    B.p1 = take(1); B.a1 = take(1); B.s1 = take(1); B.b1 = take(1); B.gA[1] = take(1); B.gB[1] = take(1); B.gC[1] = take(1);
    B.p2 = take(2); B.a2 = take(2); B.s2 = take(2); B.b2 = take(2); B.gA[2] = take(2); B.gB[2] = take(2); B.gC[2] = take(2);
    B.p3 = take(3); B.a3 = take(3); B.gA[3] = take(3); B.gB[3] = take(3); B.gC[3] = nullptr;
    S.g[0] = B.gA[0];
}

void net_input_dense(NetState &S, const float *in3, hipStream_t st) {
    const int64_t nv = (int64_t)S.d.nx * S.d.ny * S.d.nz;
    hipLaunchKernelGGL(k_net_input, dim3(blocks_for(nv, 256)), dim3(256), 0, st, S.d, in3, (bf16 *)S.b.x0);
    halo(S.d, S.b.x0, st);
}
void net_input_halo(NetState &S, hipStream_t st) { halo(S.d, S.b.x0, st); }

static const bf16 *wf_layer(const NetState &S, int l) { return (const bf16 *)S.wf + (size_t)l * 14 * 64 * 8; }
static const bf16 *wb_layer(const NetState &S, int l) { return (const bf16 *)S.wb + (size_t)l * 14 * 64 * 8; }

// forward from S.b.x0 (halo filled): out = channel 0 of the last layer (FP32 dense)
void net_forward(NetState &S, float *out, hipStream_t st) {
    const NetDims L0 = level(S.d, 0), L1 = level(S.d, 1), L2 = level(S.d, 2), L3 = level(S.d, 3);
    NetBuffers &B = S.b;
    conv_launch(S, L0, EPI_RELU, B.x0, wf_layer(S, 0), B.a0, nullptr, nullptr, st);
    hipLaunchKernelGGL(k_pool, transfer_grid(L1), dim3(256), 0, st, L1, (const bf16 *)B.a0, (bf16 *)B.p1);
    halo(L1, B.p1, st);
    conv_launch(S, L1, EPI_RELU, B.p1, wf_layer(S, 1), B.a1, nullptr, nullptr, st);
    hipLaunchKernelGGL(k_pool, transfer_grid(L2), dim3(256), 0, st, L2, (const bf16 *)B.a1, (bf16 *)B.p2);
    halo(L2, B.p2, st);
    conv_launch(S, L2, EPI_RELU, B.p2, wf_layer(S, 2), B.a2, nullptr, nullptr, st);
    hipLaunchKernelGGL(k_pool, transfer_grid(L3), dim3(256), 0, st, L3, (const bf16 *)B.a2, (bf16 *)B.p3);
    halo(L3, B.p3, st);
    conv_launch(S, L3, EPI_RELU, B.p3, wf_layer(S, 3), B.a3, nullptr, nullptr, st);
    hipLaunchKernelGGL(k_up_add, transfer_grid(L2), dim3(256), 0, st, L2, (const bf16 *)B.a2, (const bf16 *)B.a3, (bf16 *)B.s2);
    halo(L2, B.s2, st);
    conv_launch(S, L2, EPI_RELU, B.s2, wf_layer(S, 4), B.b2, nullptr, nullptr, st);
    hipLaunchKernelGGL(k_up_add, transfer_grid(L1), dim3(256), 0, st, L1, (const bf16 *)B.a1, (const bf16 *)B.b2, (bf16 *)B.s1);
    halo(L1, B.s1, st);
    conv_launch(S, L1, EPI_RELU, B.s1, wf_layer(S, 5), B.b1, nullptr, nullptr, st);
    hipLaunchKernelGGL(k_up_add, transfer_grid(L0), dim3(256), 0, st, L0, (const bf16 *)B.a0, (const bf16 *)B.b1, (bf16 *)B.s0);
    halo(L0, B.s0, st);
    conv_launch(S, L0, EPI_OUT, B.s0, wf_layer(S, 6), nullptr, out, nullptr, st);
}

static void wgrad(NetState &S, const NetDims &d, int l, const void *x, const void *g, hipStream_t st) {
    if (S.mfma_wgrad) {
        const int64_t tiles = (int64_t)(d.nx / CT_X) * ((d.ny + CT_Y - 1) / CT_Y) * ((d.nz + CT_Z - 1) / CT_Z);
        const int nb = (int)std::min<int64_t>(tiles, WG_BLOCKS_MAX);
        hipLaunchKernelGGL(k_wgrad_mfma, dim3(nb), dim3(256), 0, st, d, (const bf16 *)x, (const bf16 *)g, S.wpartial);
        hipLaunchKernelGGL(k_wgrad_sum, dim3(27), dim3(256), 0, st, (const float *)S.wpartial, nb, S.dw + (size_t)l * NET_WPL);
        return;
    }
    hipLaunchKernelGGL(k_wgrad, dim3(27 * S.nchunk), dim3(256), 0, st, d, (const bf16 *)x, (const bf16 *)g, S.nchunk, S.wpartial);
    hipLaunchKernelGGL(k_wgrad_sum, dim3(27), dim3(256), 0, st, (const float *)S.wpartial, S.nchunk, S.dw + (size_t)l * NET_WPL);
}

// backward from the output gradient held in S.b.gA[0] (halo filled): weight gradients into S.dw,
// and the input gradient's 3 channels into din (FP32 dense) when din != nullptr
void net_backward(NetState &S, float *din, hipStream_t st) {
    const NetDims L0 = level(S.d, 0), L1 = level(S.d, 1), L2 = level(S.d, 2), L3 = level(S.d, 3);
    NetBuffers &B = S.b;
    // layer 6 (level 0): g6 in gA[0]
    wgrad(S, L0, 6, B.s0, B.gA[0], st);
    conv_launch(S, L0, EPI_LIN, B.gA[0], wb_layer(S, 6), B.gB[0], nullptr, nullptr, st);        // gs0
    // layer 5 (level 1)
    hipLaunchKernelGGL(k_down_mask, transfer_grid(L1), dim3(256), 0, st, L1, (const bf16 *)B.gB[0], (const bf16 *)B.b1, (bf16 *)B.gA[1]);
    halo(L1, B.gA[1], st);
    wgrad(S, L1, 5, B.s1, B.gA[1], st);
    conv_launch(S, L1, EPI_LIN, B.gA[1], wb_layer(S, 5), B.gB[1], nullptr, nullptr, st);        // gs1
    // layer 4 (level 2)
    hipLaunchKernelGGL(k_down_mask, transfer_grid(L2), dim3(256), 0, st, L2, (const bf16 *)B.gB[1], (const bf16 *)B.b2, (bf16 *)B.gA[2]);
    halo(L2, B.gA[2], st);
    wgrad(S, L2, 4, B.s2, B.gA[2], st);
    conv_launch(S, L2, EPI_LIN, B.gA[2], wb_layer(S, 4), B.gB[2], nullptr, nullptr, st);        // gs2
    // layer 3 (level 3)
    hipLaunchKernelGGL(k_down_mask, transfer_grid(L3), dim3(256), 0, st, L3, (const bf16 *)B.gB[2], (const bf16 *)B.a3, (bf16 *)B.gA[3]);
    halo(L3, B.gA[3], st);
    wgrad(S, L3, 3, B.p3, B.gA[3], st);
    conv_launch(S, L3, EPI_LIN, B.gA[3], wb_layer(S, 3), B.gB[3], nullptr, nullptr, st);        // gp3
    // layer 2 (level 2): a2 feeds the skip sum s2 and the pool p3
    hipLaunchKernelGGL(k_up_combine, transfer_grid(L2), dim3(256), 0, st, L2, (const bf16 *)B.gB[2], (const bf16 *)B.gB[3],
                       (const bf16 *)B.a2, (bf16 *)B.gC[2]);
    halo(L2, B.gC[2], st);
    wgrad(S, L2, 2, B.p2, B.gC[2], st);
    conv_launch(S, L2, EPI_LIN, B.gC[2], wb_layer(S, 2), B.gA[2], nullptr, nullptr, st);        // gp2
    // layer 1 (level 1)
    hipLaunchKernelGGL(k_up_combine, transfer_grid(L1), dim3(256), 0, st, L1, (const bf16 *)B.gB[1], (const bf16 *)B.gA[2],
                       (const bf16 *)B.a1, (bf16 *)B.gC[1]);
    halo(L1, B.gC[1], st);
    wgrad(S, L1, 1, B.p1, B.gC[1], st);
    conv_launch(S, L1, EPI_LIN, B.gC[1], wb_layer(S, 1), B.gA[1], nullptr, nullptr, st);        // gp1
    // layer 0 (level 0)
    hipLaunchKernelGGL(k_up_combine, transfer_grid(L0), dim3(256), 0, st, L0, (const bf16 *)B.gB[0], (const bf16 *)B.gA[1],
                       (const bf16 *)B.a0, (bf16 *)B.gA[0]);
    halo(L0, B.gA[0], st);
    wgrad(S, L0, 0, B.x0, B.gA[0], st);
    if (din) conv_launch(S, L0, EPI_IN, B.gA[0], wb_layer(S, 0), nullptr, din, nullptr, st);
}

// loss of out against target over the mask; writes the output gradient into S.b.gA[0]
void net_loss(NetState &S, const float *out, const float *target, const float *mask, float inv_n, hipStream_t st) {
    hipLaunchKernelGGL(k_loss, dim3(NET_RED_BLOCKS), dim3(256), 0, st, S.d, out, target, mask, inv_n, (bf16 *)S.b.gA[0], S.lpartial);
    hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(256), 0, st, (const double *)S.lpartial, S.lsum);
    halo(S.d, S.b.gA[0], st);
}

void net_output_grad(NetState &S, const float *gd, hipStream_t st) {
    const int64_t nv = (int64_t)S.d.nx * S.d.ny * S.d.nz;
    hipLaunchKernelGGL(k_out_grad, dim3(blocks_for(nv, 256)), dim3(256), 0, st, S.d, gd, (bf16 *)S.b.gA[0]);
    halo(S.d, S.b.gA[0], st);
}

void net_adam(NetState &S, float lr, hipStream_t st) {
    S.adam_step += 1;
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(NET_LAYERS * NET_WPL, 256)), dim3(256), 0, st, S.w, S.m, S.v, (const float *)S.dw,
                       S.adam_step, lr, 0.9f, 0.999f, 1e-8f);
    net_pack(S, st);
}

}  // namespace om
#endif
