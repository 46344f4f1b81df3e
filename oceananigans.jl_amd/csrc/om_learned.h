// om_learned.h — device state of the learned preconditioner network (om_learned.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace om {

constexpr int NET_LAYERS = 7;           // convolutions of the four-level hierarchy
constexpr int NET_LEVELS = 4;
constexpr int NET_WPL = 27 * 16 * 16;   // weights per layer, FP32 master layout [tap][cin][cout]
constexpr int NET_RED_BLOCKS = 1024;    // partial sums of the loss reduction

struct NetDims {
    int nx, ny, nz;   // nx a multiple of 16; x, y periodic, z zero-padded
};

// The tensors of one application, BF16 channel-last padded boxes (one halo cell per side) of the
// level they live on (level l: the grid coarsened 2^l times in x, y and z):
//   forward  x0 a0 s0 | p1 a1 s1 b1 | p2 a2 s2 b2 | p3 a3       (x0: input; p: pooled; s: skip sums)
//   backward gA gB at every level, gC at levels 1 and 2            (back-propagated gradients)
struct NetBuffers {
    void *x0, *a0, *s0, *p1, *a1, *s1, *b1, *p2, *a2, *s2, *b2, *p3, *a3;
    void *gA[NET_LEVELS], *gB[NET_LEVELS], *gC[NET_LEVELS];
};

// Every pointer is device memory owned by the model that holds the NetState.
struct NetState {
    NetDims d{0, 0, 0};                                             // level 0 (the grid)
    NetBuffers b{};
    void *g[1] = {nullptr};                                         // = b.gA[0]: the output gradient
    float *w = nullptr, *m = nullptr, *v = nullptr, *dw = nullptr;  // FP32 master weights, Adam moments, gradients
    void *wf = nullptr, *wb = nullptr;                              // BF16 MFMA fragments: forward, adjoint
    float *out = nullptr;                                           // FP32 dense network output
    float *wpartial = nullptr;                                      // weight-gradient partial sums [max(nchunk, 512)][27][256]
    double *lpartial = nullptr;                                     // loss partial sums [NET_RED_BLOCKS]
    double *lsum = nullptr;                                         // Σ over the active cells of (out − target)²
    int nchunk = 256;
    int adam_step = 0;
    int tiled_conv = 1;   // convolutions with the input staged in LDS (1) or read per tap (0)
    int mfma_wgrad = 1;   // weight gradients on the matrix cores (1) or as FP32 sums (0)
};

// the grid sizes the network accepts: four levels of 2× coarsening with 16-voxel rows on the coarsest
inline bool net_dims_ok(int nx, int ny, int nz) { return nx % 128 == 0 && ny % 8 == 0 && nz % 8 == 0; }
// bytes of one padded activation at level l
inline size_t net_level_bytes(const NetDims &d, int l) {
    return (size_t)((d.nx >> l) + 2) * ((d.ny >> l) + 2) * ((d.nz >> l) + 2) * 16 * 2;
}
// points every buffer of S.b into `arena` (net_arena_bytes(S.d) bytes, 256-byte aligned pieces)
size_t net_arena_bytes(const NetDims &d);
void net_place_buffers(NetState &S, void *arena);

void net_pack(NetState &S, hipStream_t st);
void net_input_dense(NetState &S, const float *in3, hipStream_t st);
void net_input_halo(NetState &S, hipStream_t st);   // halos of the input box, its interior written by the caller
void net_forward(NetState &S, float *out, hipStream_t st);
void net_backward(NetState &S, float *din, hipStream_t st);
void net_loss(NetState &S, const float *out, const float *target, const float *mask, float inv_n, hipStream_t st);
void net_adam(NetState &S, float lr, hipStream_t st);
void net_output_grad(NetState &S, const float *gdense, hipStream_t st);

}  // namespace om
