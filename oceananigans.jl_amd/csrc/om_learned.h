// om_learned.h — device state of the learned preconditioner network (om_learned.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace om {

constexpr int NET_LAYERS = 4;
constexpr int NET_WPL = 27 * 16 * 16;   // weights per layer, FP32 master layout [tap][cin][cout]
constexpr int NET_RED_BLOCKS = 1024;    // partial sums of the loss reduction

struct NetDims {
    int nx, ny, nz;   // nx a multiple of 16; x, y periodic, z zero-padded
};

// Every pointer is device memory owned by the model that holds the NetState.
struct NetState {
    NetDims d{0, 0, 0};
    void *act[NET_LAYERS] = {nullptr, nullptr, nullptr, nullptr};   // BF16 padded activations: input, relu(1..3)
    void *g[2] = {nullptr, nullptr};                                // BF16 padded back-propagated gradients
    float *w = nullptr, *m = nullptr, *v = nullptr, *dw = nullptr;  // FP32 master weights, Adam moments, gradients
    void *wf = nullptr, *wb = nullptr;                              // BF16 MFMA fragments: forward, adjoint
    float *out = nullptr;                                           // FP32 dense network output
    float *wpartial = nullptr;                                      // weight-gradient partial sums [nchunk][27][256]
    double *lpartial = nullptr;                                     // loss partial sums [NET_RED_BLOCKS]
    double *lsum = nullptr;                                         // Σ over the active cells of (out − target)²
    int nchunk = 256;
    int adam_step = 0;
    int tiled_conv = 1;   // convolutions with the input staged in LDS (1) or read per tap (0)
};

void net_pack(NetState &S, hipStream_t st);
void net_input_dense(NetState &S, const float *in3, hipStream_t st);
void net_forward(NetState &S, float *out, hipStream_t st);
void net_backward(NetState &S, float *din, hipStream_t st);
void net_loss(NetState &S, const float *out, const float *target, const float *mask, float inv_n, hipStream_t st);
void net_adam(NetState &S, float lr, hipStream_t st);
void net_output_grad(NetState &S, const float *gdense, hipStream_t st);

}  // namespace om
