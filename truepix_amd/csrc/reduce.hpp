// reduce.hpp - wave64 / workgroup / grid reductions of Fp2 accumulators.
//
// Fp2 addition of canonical values is exact and associative, so every
// reduction order yields the same bits; the order here is chosen for CDNA4:
// xor-shuffles inside each wave64, one LDS slot per wave, then a grid-wide
// "last workgroup finishes" step so a round sum needs one launch and no
// separate reduce kernel.
#pragma once
#include "fp61.hpp"
#include "runtime.hpp"

namespace tp {

template <int N>
__device__ __forceinline__ void wave_sum(fp2 (&v)[N])
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const u64 r = __shfl_xor(v[k].re, off, 64);
            const u64 i = __shfl_xor(v[k].im, off, 64);
            v[k] = add2(v[k], mk2(r, i));
        }
    }
}

// Sums v over the workgroup (blockDim.x == kBlock); the total lands in thread 0.
template <int N>
__device__ __forceinline__ void block_sum(fp2 (&v)[N])
{
    __shared__ fp2 sh[N][kBlock / 64];
    wave_sum<N>(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) sh[k][wid] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            fp2 s = sh[k][0];
            for (int w = 1; w < kBlock / 64; ++w) s = add2(s, sh[k][w]);
            v[k] = s;
        }
    }
}

// Grid reduction in one launch.  Every workgroup stores its sum in
// partials[block * N + k]; the workgroup that arrives last (agent-scope
// fences around a device-wide counter make every XCD's stores visible) sums
// all partials, writes result[0..N), and re-arms the counter for the next
// launch.  Returns true in the thread that wrote the result.
template <int N>
__device__ __forceinline__ bool grid_sum(fp2 (&v)[N], fp2* partials, unsigned* counter,
                                         fp2* result)
{
    __shared__ bool last;
    const unsigned nblk = gridDim.x * gridDim.y * gridDim.z;
    const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    block_sum<N>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) partials[(size_t)bid * N + k] = v[k];
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
        last = (prev == nblk - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    fp2 acc[N];
#pragma unroll
    for (int k = 0; k < N; ++k) acc[k] = zero2();
    for (unsigned b = threadIdx.x; b < nblk; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < N; ++k) acc[k] = add2(acc[k], partials[(size_t)b * N + k]);
    }
    block_sum<N>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) result[k] = acc[k];
        atomicExch(counter, 0u);
        return true;
    }
    return false;
}

}  // namespace tp
