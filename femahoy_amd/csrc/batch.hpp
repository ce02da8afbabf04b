// Launches that can run alone or batched with the same launch of other solves.
//
// A batchable kernel is a functor K (trivially copyable: pointers, sizes, scalars) with
//   static constexpr int NT;     threads per block
//   static constexpr int MINB;   blocks per CU asked of the register allocator (0: none)
//   __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const;
// whose body takes its block index and block count from (blk, nblk) instead of blockIdx.x and
// gridDim.x. run_k launches it alone (k_run: the functor as the kernel argument). Under a
// Recorder (thread_local tl_recorder) run_k launches nothing and appends the functor's bytes and
// block count instead, so that a solve's PCG iteration can be replayed later as part of a batch:
// k_run_batch runs the same functor type for several solves in one launch, each block finding its
// solve in a table of block offsets and calling that solve's functor with its own block index and
// count. Each solve's blocks therefore do exactly what its own launch would do, and a result is
// bitwise the same alone and batched.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

namespace fem {

// One recorded launch: the batched launcher of its functor type (which also identifies the type),
// the functor's bytes and its block count.
struct BatchRec {
    void (*launch)(const void* ks, const unsigned* starts, unsigned n, unsigned blocks, hipStream_t s) = nullptr;
    uint32_t size = 0, align = 0, blocks = 0;
    std::vector<uint8_t> bytes;
};

// The launches recorded while set as tl_recorder; `unbatchable` is set when a launch that is not a
// run_k one was attempted (it was not run and cannot be replayed from the record).
struct Recorder {
    std::vector<BatchRec> recs;
    bool unbatchable = false;
};
inline thread_local Recorder* tl_recorder = nullptr;

// Marks the recording unbatchable; true when a recorder is set (the caller then launches
// nothing).
inline bool recording_unbatchable() {
    if (!tl_recorder) return false;
    tl_recorder->unbatchable = true;
    return true;
}

template <class K>
__global__ void __launch_bounds__(K::NT) k_run(K k) {
    k(blockIdx.x, gridDim.x);
}
template <class K>
__global__ void __launch_bounds__(K::NT, (K::MINB > 0 ? K::MINB : 1)) k_run_minb(K k) {
    k(blockIdx.x, gridDim.x);
}

// n solves' functors ks[0..n) with block offsets starts[0..n] (starts[n] = the launch's blocks).
template <class K>
__device__ __forceinline__ void run_batch_block(const K* __restrict__ ks, const unsigned* __restrict__ starts, unsigned n) {
    const unsigned b = blockIdx.x;
    unsigned s = 0;
    while (s + 1 < n && b >= starts[s + 1]) ++s;
    ks[s](b - starts[s], starts[s + 1] - starts[s]);
}
template <class K>
__global__ void __launch_bounds__(K::NT) k_run_batch(const K* __restrict__ ks, const unsigned* __restrict__ starts, unsigned n) {
    run_batch_block(ks, starts, n);
}
template <class K>
__global__ void __launch_bounds__(K::NT, (K::MINB > 0 ? K::MINB : 1))
    k_run_batch_minb(const K* __restrict__ ks, const unsigned* __restrict__ starts, unsigned n) {
    run_batch_block(ks, starts, n);
}

template <class K>
void batch_launch(const void* ks, const unsigned* starts, unsigned n, unsigned blocks, hipStream_t s) {
    if constexpr (K::MINB > 0)
        hipLaunchKernelGGL(k_run_batch_minb<K>, dim3(blocks), dim3(K::NT), 0, s, static_cast<const K*>(ks), starts, n);
    else
        hipLaunchKernelGGL(k_run_batch<K>, dim3(blocks), dim3(K::NT), 0, s, static_cast<const K*>(ks), starts, n);
}

// Launches k on `blocks` blocks of K::NT threads, or records it under a Recorder.
template <class K>
void run_k(const K& k, unsigned blocks, hipStream_t s) {
    static_assert(std::is_trivially_copyable<K>::value, "a batchable kernel's functor is copied as bytes");
    if (blocks == 0) return;
    if (tl_recorder) {
        BatchRec r;
        r.launch = &batch_launch<K>;
        r.size = (uint32_t)sizeof(K);
        r.align = (uint32_t)alignof(K);
        r.blocks = blocks;
        r.bytes.resize(sizeof(K));
        std::memcpy(r.bytes.data(), &k, sizeof(K));
        tl_recorder->recs.push_back(std::move(r));
        return;
    }
    if constexpr (K::MINB > 0)
        hipLaunchKernelGGL(k_run_minb<K>, dim3(blocks), dim3(K::NT), 0, s, k);
    else
        hipLaunchKernelGGL(k_run<K>, dim3(blocks), dim3(K::NT), 0, s, k);
}

}  // namespace fem
