// om_learned.hip — the learned convolutional correction to the FFT preconditioner of the
// conjugate-gradient Poisson solver, and the kernels that train it.  The reference has no such
// network; it plugs in where the reference applies its FFT preconditioner
// (conjugate_gradient_poisson_solver.jl:169-211) and is used inside flexible CG.
//
// Network: four 3×3×3 convolutions, 16 → 16 channels (the first layer's input carries 3 meaningful
// channels, the last layer's output channel 0 is the correction), ReLU after the first three,
// x and y periodic, z zero-padded.  Weights and activations are BF16, products accumulate in FP32
// on the matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Layout.  Activations are channel-last BF16 in a padded box with one halo cell on each side:
// voxel (i, j, k), i ∈ [−1, Nx], … at ((k + 1)·(Ny + 2) + j + 1)·(Nx + 2) + i + 1, 16 channels
// (32 bytes) per voxel.  A wave computes 16 consecutive voxels of one x row × 16 output channels
// as D = A·B with A = weights (16 cout × 32 k) and B = activations (32 k × 16 voxels); the K = 27·16
// contraction runs over 14 MFMA steps of two taps × 16 input channels (the 28th tap is zero).
// Lane l supplies k = 8·(l / 16) + e, e < 8, i.e. tap 2s + l/32 and input channels
// 8·((l/16) mod 2) + e — one 16-byte load per step — and receives output channels
// 4·(l / 16) + e, e < 4, of voxel l mod 16.
#include <hip/hip_runtime.h>

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
// EPI_IN: channels 0..2 as FP32 dense (the input gradient).
// -------------------------------------------------------------------------------------------
enum { EPI_RELU = 0, EPI_OUT = 1, EPI_MASK = 2, EPI_IN = 3 };

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
    if (EPI == EPI_RELU || EPI == EPI_MASK) {
        bf16x4 out;
        if (EPI == EPI_MASK) {
            const bf16x4 a = *reinterpret_cast<const bf16x4 *>(act + o * 16 + cout0);
            for (int e = 0; e < 4; ++e) out[e] = (bf16)((float)a[e] > 0.f ? acc[e] : 0.f);
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
// The same convolution with the input staged in LDS.  A block of 4 waves computes a tile of
// 16 (x) × 4 (y) × 4 (z) voxels: its 18 × 6 × 6 input voxels (one-cell halo, 16 channels, 20.7 KB)
// are loaded from the padded activation once, instead of once per tap and wave, and each wave
// (one z level of the tile) reuses every weight fragment for its four 16-voxel rows.  Same
// operands, same MFMA sequence per output: the results equal k_conv's bit for bit.
// -------------------------------------------------------------------------------------------
constexpr int CT_X = 16, CT_Y = 4, CT_Z = 4;
constexpr int CT_SX = CT_X + 2, CT_SY = CT_Y + 2, CT_SZ = CT_Z + 2;
constexpr int CT_V = CT_SX * CT_SY * CT_SZ;   // staged voxels per tile

template <int EPI>
__global__ void __launch_bounds__(256) k_conv_tiled(NetDims d, const bf16 *__restrict__ x, const bf16x8 *__restrict__ wp,
                                                    bf16 *__restrict__ ybf, float *__restrict__ yf, const bf16 *__restrict__ act) {
    __shared__ bf16x8 tile[CT_V * 2];
    const int ntx = d.nx / CT_X, nty = (d.ny + CT_Y - 1) / CT_Y;
    const int b = blockIdx.x;
    const int tx = b % ntx, rest = b / ntx;
    const int ty = rest % nty, tz = rest / nty;
    const int i0 = tx * CT_X, j0 = ty * CT_Y, k0 = tz * CT_Z;
    // stage the tile: voxel (xx, yy, zz) of the tile is padded voxel (i0 − 1 + xx, j0 − 1 + yy, k0 − 1 + zz);
    // rows and levels past the padded box's last (ny, nz) are staged as zeros and never used
    for (int q = threadIdx.x; q < CT_V * 2; q += 256) {
        const int v = q >> 1, h = q & 1;
        const int xx = v % CT_SX, r = v / CT_SX;
        const int yy = r % CT_SY, zz = r / CT_SY;
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
    f32x4 acc[CT_Y];
#pragma unroll
    for (int sgm = 0; sgm < CT_Y; ++sgm) acc[sgm] = f32x4{0.f, 0.f, 0.f, 0.f};
    bf16x8 zero;
    for (int e = 0; e < 8; ++e) zero[e] = (bf16)0.0f;
#pragma unroll 2
    for (int s = 0; s < 14; ++s) {
        const int t = 2 * s + half;
        int dx = 0, dy = 0, dz = 0;
        if (t < 27) tap_offsets(t, dx, dy, dz);
        const bf16x8 A = wp[s * 64 + lane];
        const int base = ((w + 1 + dz) * CT_SY + (1 + dy)) * CT_SX + (vi + 1 + dx);
#pragma unroll
        for (int sgm = 0; sgm < CT_Y; ++sgm) {
            const bf16x8 Bv = t < 27 ? tile[(base + sgm * CT_SX) * 2 + c0h] : zero;
            acc[sgm] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, Bv, acc[sgm], 0, 0, 0);
        }
    }
    const int cout0 = (lane >> 4) << 2;
#pragma unroll
    for (int sgm = 0; sgm < CT_Y; ++sgm) {
        const int j = j0 + sgm;
        if (j >= d.ny) break;
        const int iv = i0 + vi;
        const int64_t o = pidx(d, iv, j, k);
        if (EPI == EPI_RELU || EPI == EPI_MASK) {
            bf16x4 out;
            if (EPI == EPI_MASK) {
                const bf16x4 a = *reinterpret_cast<const bf16x4 *>(act + o * 16 + cout0);
                for (int e = 0; e < 4; ++e) out[e] = (bf16)((float)a[e] > 0.f ? acc[sgm][e] : 0.f);
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

// halo of a padded activation: x, y periodic (copied from the opposite interior edge), z zero
__global__ void __launch_bounds__(256) k_net_halo(NetDims d, bf16 *a) {
    const int64_t nxy = (int64_t)(d.nx + 2) * (d.ny + 2);
    const int64_t n = nxy * (d.nz + 2);
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n * 2) return;
    const int64_t v = t >> 1;                  // two threads per voxel: 8 channels each
    const int hh = (int)(t & 1);
    const int k = (int)(v / nxy) - 1;
    const int64_t r = v % nxy;
    const int j = (int)(r / (d.nx + 2)) - 1, i = (int)(r % (d.nx + 2)) - 1;
    const bool in_x = i >= 0 && i < d.nx, in_y = j >= 0 && j < d.ny, in_z = k >= 0 && k < d.nz;
    if (in_x && in_y && in_z) return;
    bf16x8 val;
    if (!in_z) {
        for (int e = 0; e < 8; ++e) val[e] = (bf16)0.0f;
    } else {
        const int ii = (i + d.nx) % d.nx, jj = (j + d.ny) % d.ny;
        val = *reinterpret_cast<const bf16x8 *>(a + pidx(d, ii, jj, k) * 16 + hh * 8);
    }
    *reinterpret_cast<bf16x8 *>(a + v * 16 + hh * 8) = val;
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
// layer 1 and the unused output channels of layer 4 stay zero
__global__ void __launch_bounds__(256) k_adam(float *w, float *m, float *v, const float *g, int step, float lr, float b1, float b2,
                                              float eps) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NET_LAYERS * NET_WPL) return;
    const int l = q / NET_WPL, r = q % NET_WPL, c = (r / 16) % 16, o = r % 16;
    if ((l == 0 && c >= 3) || (l == NET_LAYERS - 1 && o >= 1)) return;
    const float gq = g[q];
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

static void conv_launch(const NetState &S, int epi, const void *x, const void *wp, void *ybf, float *yf, const void *act, hipStream_t st) {
    const NetDims &d = S.d;
    const bf16 *X = (const bf16 *)x;
    const bf16x8 *W = (const bf16x8 *)wp;
    if (S.tiled_conv) {
        const int64_t tiles = (int64_t)(d.nx / CT_X) * ((d.ny + CT_Y - 1) / CT_Y) * ((d.nz + CT_Z - 1) / CT_Z);
        const dim3 g((unsigned)tiles), b(256);
        switch (epi) {
            case EPI_RELU: hipLaunchKernelGGL(k_conv_tiled<EPI_RELU>, g, b, 0, st, d, X, W, (bf16 *)ybf, yf, (const bf16 *)act); break;
            case EPI_OUT: hipLaunchKernelGGL(k_conv_tiled<EPI_OUT>, g, b, 0, st, d, X, W, (bf16 *)ybf, yf, (const bf16 *)act); break;
            case EPI_MASK: hipLaunchKernelGGL(k_conv_tiled<EPI_MASK>, g, b, 0, st, d, X, W, (bf16 *)ybf, yf, (const bf16 *)act); break;
            default: hipLaunchKernelGGL(k_conv_tiled<EPI_IN>, g, b, 0, st, d, X, W, (bf16 *)ybf, yf, (const bf16 *)act); break;
        }
        return;
    }
    const int64_t segs = (int64_t)(d.nx / 16) * d.ny * d.nz;
    const dim3 g(blocks_for(segs, 4)), b(256);
    switch (epi) {
        case EPI_RELU: hipLaunchKernelGGL(k_conv<EPI_RELU>, g, b, 0, st, d, X, W, (bf16 *)ybf, yf, (const bf16 *)act); break;
        case EPI_OUT: hipLaunchKernelGGL(k_conv<EPI_OUT>, g, b, 0, st, d, X, W, (bf16 *)ybf, yf, (const bf16 *)act); break;
        case EPI_MASK: hipLaunchKernelGGL(k_conv<EPI_MASK>, g, b, 0, st, d, X, W, (bf16 *)ybf, yf, (const bf16 *)act); break;
        default: hipLaunchKernelGGL(k_conv<EPI_IN>, g, b, 0, st, d, X, W, (bf16 *)ybf, yf, (const bf16 *)act); break;
    }
}

static void halo(const NetState &S, void *a, hipStream_t st) {
    const int64_t n = (int64_t)(S.d.nx + 2) * (S.d.ny + 2) * (S.d.nz + 2) * 2;
    hipLaunchKernelGGL(k_net_halo, dim3(blocks_for(n, 256)), dim3(256), 0, st, S.d, (bf16 *)a);
}

void net_input_dense(NetState &S, const float *in3, hipStream_t st) {
    const int64_t nv = (int64_t)S.d.nx * S.d.ny * S.d.nz;
    hipLaunchKernelGGL(k_net_input, dim3(blocks_for(nv, 256)), dim3(256), 0, st, S.d, in3, (bf16 *)S.act[0]);
    halo(S, S.act[0], st);
}

// forward from S.act[0] (halo filled): out = channel 0 of layer 4 (FP32 dense)
void net_forward(NetState &S, float *out, hipStream_t st) {
    const size_t lw = (size_t)14 * 64 * 8;
    for (int l = 0; l < NET_LAYERS - 1; ++l) {
        conv_launch(S, EPI_RELU, S.act[l], (const bf16 *)S.wf + l * lw, S.act[l + 1], nullptr, nullptr, st);
        halo(S, S.act[l + 1], st);
    }
    conv_launch(S, EPI_OUT, S.act[NET_LAYERS - 1], (const bf16 *)S.wf + (NET_LAYERS - 1) * lw, nullptr, out, nullptr, st);
}

// backward from the output gradient held in S.g[0] (halo filled): weight gradients into S.dw,
// and the input gradient's 3 channels into din (FP32 dense) when din != nullptr
void net_backward(NetState &S, float *din, hipStream_t st) {
    const size_t lw = (size_t)14 * 64 * 8;
    int cur = 0;
    for (int l = NET_LAYERS - 1; l >= 0; --l) {
        hipLaunchKernelGGL(k_wgrad, dim3(27 * S.nchunk), dim3(256), 0, st, S.d, (const bf16 *)S.act[l], (const bf16 *)S.g[cur],
                           S.nchunk, S.wpartial);
        hipLaunchKernelGGL(k_wgrad_sum, dim3(27), dim3(256), 0, st, (const float *)S.wpartial, S.nchunk, S.dw + (size_t)l * NET_WPL);
        if (l > 0) {
            conv_launch(S, EPI_MASK, S.g[cur], (const bf16 *)S.wb + l * lw, S.g[1 - cur], nullptr, S.act[l], st);
            halo(S, S.g[1 - cur], st);
            cur = 1 - cur;
        } else if (din) {
            conv_launch(S, EPI_IN, S.g[cur], (const bf16 *)S.wb, nullptr, din, nullptr, st);
        }
    }
}

// loss of out against target over the mask; writes the output gradient into S.g[0]
void net_loss(NetState &S, const float *out, const float *target, const float *mask, float inv_n, hipStream_t st) {
    hipLaunchKernelGGL(k_loss, dim3(NET_RED_BLOCKS), dim3(256), 0, st, S.d, out, target, mask, inv_n, (bf16 *)S.g[0], S.lpartial);
    hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(256), 0, st, (const double *)S.lpartial, S.lsum);
    halo(S, S.g[0], st);
}

void net_output_grad(NetState &S, const float *gd, hipStream_t st) {
    const int64_t nv = (int64_t)S.d.nx * S.d.ny * S.d.nz;
    hipLaunchKernelGGL(k_out_grad, dim3(blocks_for(nv, 256)), dim3(256), 0, st, S.d, gd, (bf16 *)S.g[0]);
    halo(S, S.g[0], st);
}

void net_adam(NetState &S, float lr, hipStream_t st) {
    S.adam_step += 1;
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(NET_LAYERS * NET_WPL, 256)), dim3(256), 0, st, S.w, S.m, S.v, (const float *)S.dw,
                       S.adam_step, lr, 0.9f, 0.999f, 1e-8f);
    net_pack(S, st);
}

}  // namespace om
#endif
