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
    __shared__ double sh[NV][kBlock / kWave];
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

// Deterministic grid-wide sum: every block stores its NV partial sums, the
// last block to arrive adds all partials in a fixed order and calls
// fin(totals) from thread 0.  The counter is reset for the next launch.
// Partial stores/loads and the arrival counter use agent-scope atomics so
// they are coherent across the XCDs' separate L2 caches.
template <int NV, class Fin>
__device__ __forceinline__ void grid_sum(double (&v)[NV], double* partials, unsigned* counter, Fin fin) {
    block_sum<NV>(v);
    __shared__ bool last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i)
            __hip_atomic_store(&partials[(size_t)NV * blockIdx.x + i], v[i], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    double t[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) t[i] = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
#pragma unroll
        for (int i = 0; i < NV; ++i)
            t[i] += __hip_atomic_load(&partials[(size_t)NV * b + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    block_sum<NV>(t);
    if (threadIdx.x == 0) {
        fin(t);
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Grid-wide max of one non-negative value, same scheme.
template <class Fin>
__device__ __forceinline__ void grid_max(double v, double* partials, unsigned* counter, Fin fin) {
    __shared__ double sh[kBlock / kWave];
    __shared__ bool last;
    int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    v = wave_max(v);
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int k = 0; k < (int)(blockDim.x / kWave); ++k) m = m > sh[k] ? m : sh[k];
        __hip_atomic_store(&partials[blockIdx.x], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    double m = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
        double x = __hip_atomic_load(&partials[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        m = m > x ? m : x;
    }
    m = wave_max(m);
    __syncthreads();
    if (lane == 0) sh[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0.0;
        for (int k = 0; k < (int)(blockDim.x / kWave); ++k) mm = mm > sh[k] ? mm : sh[k];
        fin(mm);
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace asend
