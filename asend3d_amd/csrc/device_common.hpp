// Device-side helpers shared by the solve-path kernels: wavefront/block
// reductions, the deterministic last-block grid reduction and the XCD-aware
// block remap for gfx950 (256 CUs in 8 XCDs, 64-lane wavefronts).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace asend {

constexpr int kWave = 64;
constexpr int kBlock = 256;  // threads per block for streaming kernels
constexpr int kNumXcd = 8;
constexpr int kMaxBlockWaves = 16;  // block reductions accept up to 1024 threads

// Element blocks of brick8 meshes by rows: record (e, a) holds the eight 3x3
// blocks K_ab (b = 0..7) of element e's row a, row-major, at 9 b of its
// kRecStride doubles (576 bytes, a multiple of 16); element e's records are
// contiguous, record (e, a) at kRecStride (8 e + a).  Record (e, a) is
// consumed by exactly one operator row (node conn[e][a]).
constexpr int kRecStride = 72;

// Logical block index such that each XCD (blocks are dispatched round-robin
// over the 8 XCDs) processes one contiguous range of logical blocks, so rows
// that share x-vector entries share an L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    int per = (nb + kNumXcd - 1) / kNumXcd;
    int xcd = b % kNumXcd, idx = b / kNumXcd;
    int full = nb % kNumXcd == 0 ? kNumXcd : nb % kNumXcd;  // XCDs owning 'per' blocks
    // XCD x owns blocks x, x+8, ...; count of blocks owned by XCDs < xcd:
    int before = (xcd < full) ? xcd * per : full * per + (xcd - full) * (per - 1);
    return before + idx;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, kWave);
        v = v > w ? v : w;
    }
    return v;
}

// Sum of NV values over the whole block; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double sh[NV][kMaxBlockWaves];
    int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) sh[i][w] = v[i];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
            for (int k = 0; k < (int)(blockDim.x / kWave); ++k) s += sh[i][k];
            v[i] = s;
        }
    }
    __syncthreads();
}

// Block-level partial sums for a two-kernel grid reduction: thread 0 of every
// block stores its NV sums at partials[NV*blockIdx.x + i]; a single-block
// finalize kernel launched next on the same stream combines them.
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* partials) {
    block_sum<NV>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) partials[(size_t)NV * blockIdx.x + i] = v[i];
}

// Block-level maximum of a non-negative value, same scheme.
__device__ __forceinline__ void block_partial_max(double v, double* partials) {
    __shared__ double sh[kMaxBlockWaves];
    v = wave_max(v);
    if (threadIdx.x % kWave == 0) sh[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int k = 0; k < (int)(blockDim.x / kWave); ++k) m = m > sh[k] ? m : sh[k];
        partials[blockIdx.x] = m;
    }
}

}  // namespace asend
