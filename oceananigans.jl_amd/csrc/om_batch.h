// om_batch.h — one kernel body, two launch forms: a launch for one model, and a launch over a
// table of members (an ensemble of independent models of any sizes) that covers every member's
// blocks in one grid.
//
// A kernel kind K is a struct with kThreads, kMinBlocks and a __device__ `run(const A &, const Blk &)`
// that reads its block coordinates from the Blk it is given, never from blockIdx/gridDim.
//   k_one<K, A>(a)          hardware block coordinates
//   k_many<K, A>(tab, n)    a flat 1-D grid; member m owns blocks [tab[m].begin, tab[m].begin +
//                           gx·gy·gz) with begin a multiple of 8, so that a member's block id keeps
//                           its residue mod 8 (the XCD the dispatcher gives it) and XCD-aware tile
//                           orders stay valid; the blocks between members exit at once.
// Host launches go through launch_k<K>(grid, block, lds, stream, args).  With a Recorder installed
// on the calling thread, the launch is appended to it instead (kind, grid, block, LDS bytes and the
// argument bytes), so an ensemble can replay every member's launch sequence as one member-table
// launch per position.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <vector>

namespace om {

struct Blk {
    uint32_t x, y, z, gx, gy, gz;
};

template <class A> struct BatchEntry {
    A a;
    uint32_t gx, gy, gz, begin;
};

template <class K, class A> __global__ void __launch_bounds__(K::kThreads, K::kMinBlocks) k_one(A a) {
    K::run(a, Blk{blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, gridDim.y, gridDim.z});
}

template <class K, class A>
__global__ void __launch_bounds__(K::kThreads, K::kMinBlocks) k_many(const BatchEntry<A> *__restrict__ tab, int n) {
    const uint32_t id = blockIdx.x;
    int lo = 0, hi = n - 1;   // the last member whose range starts at or before id
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].begin <= id) lo = mid;
        else hi = mid - 1;
    }
    const BatchEntry<A> &E = tab[lo];
    const uint32_t local = id - E.begin;
    if (local >= E.gx * E.gy * E.gz) return;
    const uint32_t x = local % E.gx, r = local / E.gx;
    K::run(E.a, Blk{x, r % E.gy, r / E.gy, E.gx, E.gy, E.gz});
}

// ---------------------------------------------------------------------------------------------
// recording
// ---------------------------------------------------------------------------------------------
struct LaunchRec {
    const void *kind;                                  // identifies (K, A): &launch_many<K, A>
    dim3 grid, block;
    size_t lds;
    std::vector<unsigned char> args;                   // the A bytes
    size_t entry_size;                                 // sizeof(BatchEntry<A>)
    void (*fill)(unsigned char *dst, const LaunchRec &r, uint32_t begin);
    void (*launch_many)(const void *tab, int n, uint32_t blocks, dim3 block, size_t lds, hipStream_t st);
};

struct Recorder {
    std::vector<LaunchRec> recs;
};

// the recorder of the calling thread (nullptr: launches run)
Recorder *&thread_recorder();

template <class K, class A> void batch_fill(unsigned char *dst, const LaunchRec &r, uint32_t begin) {
    BatchEntry<A> e;
    std::memset(&e, 0, sizeof e);
    std::memcpy(&e.a, r.args.data(), sizeof(A));
    e.gx = r.grid.x; e.gy = r.grid.y; e.gz = r.grid.z; e.begin = begin;
    std::memcpy(dst, &e, sizeof e);
}

template <class K, class A> void batch_launch_many(const void *tab, int n, uint32_t blocks, dim3 block, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL((k_many<K, A>), dim3(blocks), block, lds, st, static_cast<const BatchEntry<A> *>(tab), n);
}

template <class K, class A> void launch_k(dim3 grid, dim3 block, size_t lds, hipStream_t st, const A &a) {
    if (Recorder *R = thread_recorder()) {
        LaunchRec r;
        r.kind = reinterpret_cast<const void *>(&batch_launch_many<K, A>);
        r.grid = grid; r.block = block; r.lds = lds;
        r.args.resize(sizeof(A));
        std::memcpy(r.args.data(), &a, sizeof(A));
        r.entry_size = sizeof(BatchEntry<A>);
        r.fill = &batch_fill<K, A>;
        r.launch_many = &batch_launch_many<K, A>;
        R->recs.push_back(std::move(r));
        return;
    }
    hipLaunchKernelGGL((k_one<K, A>), grid, block, lds, st, a);
}

}  // namespace om
