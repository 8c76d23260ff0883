// reduce.hpp - wave64 / workgroup reductions of Fp2 accumulators and the
// two-launch grid reduction.
//
// Fp2 addition of canonical values is exact and associative, so every
// reduction order yields the same bits; the order here is chosen for CDNA4:
// DPP lane moves inside each wave64, one LDS slot per wave, one partial per
// workgroup, then a one-workgroup finalize launch.  Cross-workgroup
// visibility comes from the kernel boundary: an in-kernel "last workgroup"
// scheme needs agent-scope release/acquire fences, which on gfx950 write back
// and invalidate the whole XCD L2 (buffer_wbl2 / buffer_inv) in every
// workgroup and throttle a streaming kernel.
#pragma once
#include "fp61.hpp"
#include "runtime.hpp"

namespace tp {

// One DPP step of a wave64 sum: every lane of the rows in ROW_MASK adds the
// value CTRL selects (a lane move inside the VALU, no LDS round trip); lanes
// of the other rows add 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ u64 dpp_u64(u64 x)
{
    u32 lo, hi;
    if (ROW_MASK == 0xf) {   // every lane has a source: no old value to keep
        lo = (u32)__builtin_amdgcn_mov_dpp((int)(u32)x, CTRL, 0xf, 0xf, false);
        hi = (u32)__builtin_amdgcn_mov_dpp((int)(u32)(x >> 32), CTRL, 0xf, 0xf, false);
    } else {
        lo = (u32)__builtin_amdgcn_update_dpp(0, (int)(u32)x, CTRL, ROW_MASK, 0xf, false);
        hi = (u32)__builtin_amdgcn_update_dpp(0, (int)(u32)(x >> 32), CTRL, ROW_MASK, 0xf, false);
    }
    return ((u64)hi << 32) | lo;
}

template <int CTRL, int ROW_MASK, int W>
__device__ __forceinline__ void dpp_add_step(u64 (&w)[W])
{
#pragma unroll
    for (int k = 0; k < W; ++k) w[k] += dpp_u64<CTRL, ROW_MASK>(w[k]);
}

template <int W>
__device__ __forceinline__ void fold_words(u64 (&w)[W])
{
#pragma unroll
    for (int k = 0; k < W; ++k) w[k] = fold1(w[k]);
}

// Sum over the 64 lanes of a wave (all of them active), the total in every
// lane.  DPP steps: xor 1 and xor 2 inside each quad (quad_perm), the
// 8-lane and 16-lane mirrors inside each row (after which every lane of a
// row holds the row's sum), then row_bcast:15 into rows 1 and 3 and
// row_bcast:31 into rows 2 and 3, leaving the total in lane 63, which
// readlane hands to every lane.  The words add as plain u64 with a fold
// after every second step: from canonical inputs a sum of four values of at
// most 2^61 + 6 stays below 2^64, and one reduction at the end gives the
// canonical total, the same bits as any order of canonical additions.
template <int N>
__device__ __forceinline__ void wave_sum(fp2 (&v)[N])
{
    u64 w[2 * N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        w[2 * k] = v[k].re;
        w[2 * k + 1] = v[k].im;
    }
    dpp_add_step<0xb1, 0xf>(w);    // quad_perm [1,0,3,2]
    dpp_add_step<0x4e, 0xf>(w);    // quad_perm [2,3,0,1]
    fold_words(w);
    dpp_add_step<0x141, 0xf>(w);   // row_half_mirror
    dpp_add_step<0x140, 0xf>(w);   // row_mirror
    fold_words(w);
    dpp_add_step<0x142, 0xa>(w);   // row_bcast:15, rows 1 and 3
    dpp_add_step<0x143, 0xc>(w);   // row_bcast:31, rows 2 and 3
#pragma unroll
    for (int k = 0; k < 2 * N; ++k) {
        const u64 c = canon(fold1(w[k]));
        const u32 lo = (u32)__builtin_amdgcn_readlane((int)(u32)c, 63);
        const u32 hi = (u32)__builtin_amdgcn_readlane((int)(u32)(c >> 32), 63);
        w[k] = ((u64)hi << 32) | lo;
    }
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = mk2(w[2 * k], w[2 * k + 1]);
}

// wave_sum when only lanes 0 .. nlanes - 1 hold nonzero values (nlanes the
// same in every lane): the DPP steps that would only add zeros are skipped.
// After the quad steps lane 0 holds the sum of lanes 0-3, after the two
// mirrors of lanes 0-15, after row_bcast:15 lane 31 holds lanes 0-31, and
// after row_bcast:31 lane 63 holds all; the total goes to every lane.
template <int N>
__device__ __forceinline__ void wave_sum_upto(fp2 (&v)[N], unsigned nlanes)
{
    if (nlanes > 32) {
        wave_sum<N>(v);
        return;
    }
    u64 w[2 * N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        w[2 * k] = v[k].re;
        w[2 * k + 1] = v[k].im;
    }
    if (nlanes > 1) dpp_add_step<0xb1, 0xf>(w);    // quad_perm [1,0,3,2]
    if (nlanes > 2) dpp_add_step<0x4e, 0xf>(w);    // quad_perm [2,3,0,1]
    fold_words(w);
    if (nlanes > 4) dpp_add_step<0x141, 0xf>(w);   // row_half_mirror
    if (nlanes > 8) dpp_add_step<0x140, 0xf>(w);   // row_mirror
    fold_words(w);
    const int src = nlanes > 16 ? 31 : 0;
    if (nlanes > 16) dpp_add_step<0x142, 0xa>(w);  // row_bcast:15, rows 1 and 3
#pragma unroll
    for (int k = 0; k < 2 * N; ++k) {
        const u64 c = canon(fold1(w[k]));
        const u32 lo = (u32)__builtin_amdgcn_readlane((int)(u32)c, src);
        const u32 hi = (u32)__builtin_amdgcn_readlane((int)(u32)(c >> 32), src);
        w[k] = ((u64)hi << 32) | lo;
    }
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = mk2(w[2 * k], w[2 * k + 1]);
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
