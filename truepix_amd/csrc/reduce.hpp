// reduce.hpp - wave64 / workgroup reductions of Fp2 accumulators and the
// two-launch grid reduction.
//
// Fp2 addition of canonical values is exact and associative, so every
// reduction order yields the same bits; the order here is chosen for CDNA4:
// xor-shuffles inside each wave64, one LDS slot per wave, one partial per
// workgroup, then a one-workgroup finalize launch.  Cross-workgroup
// visibility comes from the kernel boundary: an in-kernel "last workgroup"
// scheme needs agent-scope release/acquire fences, which on gfx950 write back
// and invalidate the whole XCD L2 (buffer_wbl2 / buffer_inv) in every
// workgroup and throttle a streaming kernel.
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

// Sums v over the workgroup (blockDim.x a multiple of 64, at most 1024); the
// total lands in thread 0.
template <int N>
__device__ __forceinline__ void block_sum(fp2 (&v)[N])
{
    __shared__ fp2 sh[N][16];
    wave_sum<N>(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) sh[k][wid] = v[k];
    }
    __syncthreads();
    if (nw <= 4) {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < N; ++k) {
                fp2 s = sh[k][0];
                for (int w = 1; w < nw; ++w) s = add2(s, sh[k][w]);
                v[k] = s;
            }
        }
    } else if (wid == 0) {
        // up to 16 wave totals: lanes 0..15 of wave 0, four xor-shuffle levels
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = lane < nw ? sh[k][lane] : zero2();
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) {
#pragma unroll
            for (int k = 0; k < N; ++k) {
                const u64 r = __shfl_xor(v[k].re, off, 64);
                const u64 i = __shfl_xor(v[k].im, off, 64);
                v[k] = add2(v[k], mk2(r, i));
            }
        }
    }
}

__device__ __forceinline__ unsigned linear_block_id()
{
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}

// First half of a grid reduction: the workgroup total goes to
// partials[block * N + k].
template <int N>
__device__ __forceinline__ void block_partial(fp2 (&v)[N], fp2* partials)
{
    block_sum<N>(v);
    if (threadIdx.x == 0) {
        const size_t b = linear_block_id();
#pragma unroll
        for (int k = 0; k < N; ++k) partials[b * N + k] = v[k];
    }
}

// Sums n_blocks partials of N accumulators inside one workgroup; the total
// lands in thread 0's acc.
template <int N>
__device__ __forceinline__ void sum_partials(const fp2* partials, unsigned n_blocks, fp2 (&acc)[N])
{
#pragma unroll
    for (int k = 0; k < N; ++k) acc[k] = zero2();
    for (unsigned b = threadIdx.x; b < n_blocks; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < N; ++k) acc[k] = add2(acc[k], partials[(size_t)b * N + k]);
    }
    block_sum<N>(acc);
}

// Second half: one workgroup of kBlock threads writes result[0..N).
template <int N>
__global__ void __launch_bounds__(kBlock) k_finalize_sum(const fp2* partials, unsigned n_blocks,
                                                         fp2* result)
{
    fp2 acc[N];
    sum_partials<N>(partials, n_blocks, acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) result[k] = acc[k];
    }
}

// One workgroup sums n_blocks partials and publishes the N totals to mapped
// host memory, then the sequence number (MappedSlot protocol).
template <int N>
__global__ void __launch_bounds__(kBlock) k_publish_sum(const fp2* partials, unsigned n_blocks,
                                                        fp2* host_vals, volatile unsigned* host_seq,
                                                        unsigned seq)
{
    fp2 acc[N];
    sum_partials<N>(partials, n_blocks, acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) host_vals[k] = acc[k];
        __threadfence_system();
        *host_seq = seq;
    }
}

// k_publish_sum with 1024 threads: for many partials (a few per thread).
template <int N>
__global__ void __launch_bounds__(1024) k_publish_wide(const fp2* partials, unsigned n_blocks,
                                                       fp2* host_vals, volatile unsigned* host_seq,
                                                       unsigned seq)
{
    fp2 acc[N];
    sum_partials<N>(partials, n_blocks, acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) host_vals[k] = acc[k];
        __threadfence_system();
        *host_seq = seq;
    }
}

}  // namespace tp
