// native_fp2.hip - the drop-in C ABI of the reference's native_fp2 library
// (TP/native_fp2/native_fp2.cu:824-1022, TP/native_fp2/native_fp2.cpp:11-160),
// implemented with CDNA4 kernels.  Same 14 symbols, argument meaning,
// SoA (re[], im[]) layout, synchronous semantics and session life cycle.
//
// Differences from the reference, all deliberate:
//  * every size runs on the GPU; there is no host fallback.  A device failure
//    is reported through tp_last_error() (void entry points) or a negative
//    return (int entry points) instead of silently recomputing on the host;
//  * the early-round claim sums all four evaluation points in one launch
//    (per-workgroup partials, then a one-workgroup finalize launch);
//  * beta and the input MLE are built from LDS-resident half tables in one
//    launch instead of n_z doubling launches.
#include "claim_terms.hpp"
#include "fp61.hpp"
#include "kernels.hpp"
#include "layer_bodies.hpp"
#include "reduce.hpp"
#include "runtime.hpp"

#include <stdint.h>
#include <string.h>

using namespace tp;

namespace tp {
namespace kern {

// Early-round claim over resident [wire][nc] tables (session form,
// TP/native_fp2/native_fp2_core.hpp:200-248), over the layer's claim terms
// (body::early_claim): thread (pair p, term chunk) sums its terms at
// t = 0, 1, -1, 2, scales by beta_t(p) and joins the grid sum.
__global__ void k_early_claim(int n_gates, int gpt, int n_chunks, const int* __restrict__ gt,
                              const int* __restrict__ in0, const int* __restrict__ in1,
                              const u64* __restrict__ z1r, const u64* __restrict__ z1i, int nc,
                              const u64* __restrict__ vr, const u64* __restrict__ vi,
                              const u64* __restrict__ br, const u64* __restrict__ bi,
                              fp2* partials)
{
    body::early_claim(n_gates, gpt, n_chunks, gt, in0, in1, z1r, z1i, nc, vr, vi, br, bi, partials,
                      blockIdx.x, blockIdx.x);
}

// Claim-term coefficients from a frame's z1 values (body::term_coeffs: one
// wave per term, so n_terms * 64 threads).
__global__ void k_term_coeffs(const u64* __restrict__ z1r, const u64* __restrict__ z1i,
                              const int* __restrict__ off, const int* __restrict__ src,
                              const int* __restrict__ neg, int n_terms, u64* __restrict__ tzr,
                              u64* __restrict__ tzi)
{
    body::term_coeffs(z1r, z1i, off, src, neg, n_terms, tzr, tzi, blockIdx.x);
}

// Fold every wire row [n_wires][n_in] -> [n_wires][n_in/2] and beta in one
// launch (TP/native_fp2/native_fp2_core.hpp:250-279).
__global__ void k_fold_rows(const u64* __restrict__ ir, const u64* __restrict__ ii, int n_wires,
                            int log_out, u64 rr, u64 ri, u64* __restrict__ orr,
                            u64* __restrict__ oi, const u64* __restrict__ bir,
                            const u64* __restrict__ bii, u64* __restrict__ bor,
                            u64* __restrict__ boi)
{
    body::fold_rows(ir, ii, n_wires, log_out, mk2(rr, ri), orr, oi, bir, bii, bor, boi, blockIdx.x,
                    gridDim.x);
}

}  // namespace kern

ClaimShape claim_shape(int n_gates, int n_pairs, long long target)
{
    ClaimShape c;
    long long chunks = target / (n_pairs > 0 ? n_pairs : 1);
    if (chunks < 1) chunks = 1;
    if (chunks > n_gates) chunks = n_gates;
    c.gpt = (int)((n_gates + chunks - 1) / chunks);
    if (c.gpt < 1) c.gpt = 1;
    c.n_chunks = (n_gates + c.gpt - 1) / c.gpt;
    const long long items = (long long)c.n_chunks * n_pairs;
    c.blocks = (unsigned)((items + kBlock - 1) / kBlock);
    if (c.blocks < 1) c.blocks = 1;
    return c;
}

int gates_per_thread(int n_gates, int n_pairs)
{
    const long long target = 1LL << 16;
    long long chunks = target / (n_pairs > 0 ? n_pairs : 1);
    if (chunks < 1) chunks = 1;
    if (chunks > n_gates) chunks = n_gates;
    long long gpt = (n_gates + chunks - 1) / chunks;
    return (int)(gpt < 1 ? 1 : gpt);
}

}  // namespace tp

namespace {

// ---------------------------------------------------------------- kernels --

__global__ void k_fold(const u64* __restrict__ ir, const u64* __restrict__ ii, size_t n_out,
                       u64 rr, u64 ri, u64* __restrict__ orr, u64* __restrict__ oi)
{
    const fp2 r = mk2(rr, ri);
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n_out;
         k += (size_t)gridDim.x * blockDim.x) {
        const fp2 v = fold2(mk2(ir[2 * k], ii[2 * k]), mk2(ir[2 * k + 1], ii[2 * k + 1]), r);
        orr[k] = v.re;
        oi[k] = v.im;
    }
}

__global__ void k_expand(const u64* __restrict__ ir, const u64* __restrict__ ii, size_t n_src,
                         size_t nc, u64* __restrict__ orr, u64* __restrict__ oi)
{
    const size_t total = n_src * nc;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < total;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t s = k / nc;
        orr[k] = ir[s];
        oi[k] = ii[s];
    }
}

// Non-session early claim with explicit -1 / 2 tables
// (TP/native_fp2/native_fp2_core.hpp:82-137); fact tables are [wire][nc/2].
__global__ void k_sumcheck_early(int n_gates, int gpt, const int* __restrict__ gt,
                                 const int* __restrict__ in0, const int* __restrict__ in1,
                                 const u64* __restrict__ z1r, const u64* __restrict__ z1i, int nc,
                                 const u64* __restrict__ vr, const u64* __restrict__ vi,
                                 const u64* __restrict__ f0r, const u64* __restrict__ f0i,
                                 const u64* __restrict__ f1r, const u64* __restrict__ f1i,
                                 const u64* __restrict__ br, const u64* __restrict__ bi,
                                 const u64* __restrict__ bf0r, const u64* __restrict__ bf0i,
                                 const u64* __restrict__ bf1r, const u64* __restrict__ bf1i,
                                 fp2* partials)
{
    const int hn = nc >> 1;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int g0 = blockIdx.y * gpt;
    const int g1 = min(n_gates, g0 + gpt);
    fp2 acc[4] = {zero2(), zero2(), zero2(), zero2()};
    if (p < hn) {
        const size_t c = 2 * (size_t)p, h = (size_t)p;
        fp2 v0 = zero2(), v1 = zero2(), v2 = zero2(), v3 = zero2();
        for (int g = g0; g < g1; ++g) {
            const int tpg = gt[g];
            const size_t wa = (size_t)in0[g], wb = (size_t)in1[g];
            const fp2 z = mk2(z1r[g], z1i[g]);
            const size_t a = wa * nc + c, b = wb * nc + c;
            const size_t fa = wa * hn + h, fb = wb * hn + h;
            v0 = add2(v0, mul2(gatefn(tpg, mk2(vr[a], vi[a]), mk2(vr[b], vi[b])), z));
            v1 = add2(v1, mul2(gatefn(tpg, mk2(vr[a + 1], vi[a + 1]), mk2(vr[b + 1], vi[b + 1])), z));
            v2 = add2(v2, mul2(gatefn(tpg, mk2(f0r[fa], f0i[fa]), mk2(f0r[fb], f0i[fb])), z));
            v3 = add2(v3, mul2(gatefn(tpg, mk2(f1r[fa], f1i[fa]), mk2(f1r[fb], f1i[fb])), z));
        }
        acc[0] = mul2(v0, mk2(br[c], bi[c]));
        acc[1] = mul2(v1, mk2(br[c + 1], bi[c + 1]));
        acc[2] = mul2(v2, mk2(bf0r[h], bf0i[h]));
        acc[3] = mul2(v3, mk2(bf1r[h], bf1i[h]));
    }
    block_partial<4>(acc, partials);
}

// Low-bit half of an eq/beta table in LDS: tab[t] = init * prod_{j<lb}
// (bit_j(t) ? z_j : 1 - z_j) for t < 2^lb, by in-place doubling.
__device__ void lds_half_beta(fp2* tab, const u64* zr, const u64* zi, int lb, fp2 init)
{
    if (threadIdx.x == 0) tab[0] = init;
    __syncthreads();
    for (int j = 0; j < lb; ++j) {
        const int cur = 1 << j;
        const fp2 z = mk2(zr[j], zi[j]);
        const fp2 omz = sub2(one2(), z);
        for (int t = threadIdx.x; t < cur; t += blockDim.x) {
            const fp2 b = tab[t];
            tab[t + cur] = mul2(b, z);
            tab[t] = mul2(b, omz);
        }
        __syncthreads();
    }
}

// Product of the high-bit factors for row index `row` (bits j >= lb of the
// element index), LSB-first.
__device__ fp2 high_factor(size_t row, const u64* zr, const u64* zi, int lb, int n_z)
{
    fp2 h = one2();
    for (int j = lb; j < n_z; ++j) {
        const fp2 z = mk2(zr[j], zi[j]);
        h = mul2(h, ((row >> (j - lb)) & 1) ? z : sub2(one2(), z));
    }
    return h;
}

// beta table (TP/native_fp2/native_fp2_core.hpp:139-159): each 256-thread
// workgroup writes 256 consecutive entries = low table (LDS) x its row factor.
__global__ void k_beta(const u64* __restrict__ zr, const u64* __restrict__ zi, int n_z,
                       u64 init_r, u64 init_i, u64* __restrict__ orr, u64* __restrict__ oi)
{
    __shared__ fp2 lo[kBlock];
    __shared__ fp2 hi;
    const int lb = n_z < 8 ? n_z : 8;
    lds_half_beta(lo, zr, zi, lb, mk2(init_r, init_i));
    if (threadIdx.x == 0) hi = high_factor(blockIdx.x, zr, zi, lb, n_z);
    __syncthreads();
    const int t = threadIdx.x;
    if (t < (1 << lb)) {
        const fp2 v = mul2(lo[t], hi);
        const size_t k = ((size_t)blockIdx.x << lb) + t;
        orr[k] = v.re;
        oi[k] = v.im;
    }
}

// Input MLE at z for a base-field vector of 2^n_z entries
// (TP/native_fp2/native_fp2_core.hpp:300-364): sum_x beta_z(x) * from_base(in[x]).
// Rows of 2^lb inputs share the LDS low table; workgroups stride over rows.
__global__ void k_mle_base(const u64* __restrict__ in, const u64* __restrict__ zr,
                           const u64* __restrict__ zi, int n_z, fp2* partials)
{
    __shared__ fp2 lo[1024];
    __shared__ fp2 hrow;
    const int lb = n_z < 10 ? n_z : 10;
    const int row_len = 1 << lb;
    const size_t n_rows = (size_t)1 << (n_z - lb);
    lds_half_beta(lo, zr, zi, lb, one2());
    fp2 acc[1] = {zero2()};
    for (size_t row = blockIdx.x; row < n_rows; row += gridDim.x) {
        if (threadIdx.x == 0) hrow = high_factor(row, zr, zi, lb, n_z);
        __syncthreads();
        fp2 dot = zero2();
        const u64* src = in + row * row_len;
        for (int t = threadIdx.x; t < row_len; t += blockDim.x)
            dot = add2(dot, smul2(fred64(src[t]), lo[t]));
        acc[0] = add2(acc[0], mul2(dot, hrow));
        __syncthreads();
    }
    block_partial<1>(acc, partials);
}

// ------------------------------------------------------------ host state --

struct Pool {
    DevBuf a_r, a_i, b_r, b_i;                  // fold / expand in-out
    DevBuf gt, i0, i1, z1r, z1i;                // gate tables
    DevBuf v_r, v_i, f0_r, f0_i, f1_r, f1_i;    // non-session sumcheck tables
    DevBuf be_r, be_i, bf0_r, bf0_i, bf1_r, bf1_i;
    DevBuf red;                                 // result + partials
    DevBuf mle_in, z_r, z_i;
};
Pool g_pool;

// Reduction scratch: [result (256 B)] [partials ...].
struct RedView { fp2* result; fp2* partials; };
bool red_view(DevBuf& buf, size_t n_blocks, int n_acc, RedView& v)
{
    if (!buf.ensure(256 + n_blocks * n_acc * sizeof(fp2))) return false;
    char* base = buf.as<char>();
    v.result = reinterpret_cast<fp2*>(base);
    v.partials = reinterpret_cast<fp2*>(base + 256);
    return true;
}

// Finalize launch of a grid reduction with N accumulators.
template <int N>
bool finalize(const RedView& rv, size_t n_blocks)
{
    k_finalize_sum<N><<<1, kBlock, 0, stream()>>>(rv.partials, (unsigned)n_blocks, rv.result);
    TP_HIP_OK(hipGetLastError());
    return true;
}

bool h2d(DevBuf& b, const void* src, size_t bytes)
{
    if (!b.ensure(bytes)) return false;
    if (bytes) TP_HIP_OK(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, stream()));
    return true;
}

bool launch_ok() { TP_HIP_OK(hipGetLastError()); return true; }

bool gpu_fold(const u64* in_r, const u64* in_i, size_t n_out, u64 rr, u64 ri, u64* out_r,
              u64* out_i)
{
    Pool& P_ = g_pool;
    if (!h2d(P_.a_r, in_r, 2 * n_out * 8) || !h2d(P_.a_i, in_i, 2 * n_out * 8)) return false;
    if (!P_.b_r.ensure(n_out * 8) || !P_.b_i.ensure(n_out * 8)) return false;
    k_fold<<<grid_for(n_out), kBlock, 0, stream()>>>(P_.a_r.as<u64>(), P_.a_i.as<u64>(), n_out,
                                                      rr, ri, P_.b_r.as<u64>(), P_.b_i.as<u64>());
    if (!launch_ok()) return false;
    TP_HIP_OK(hipMemcpyAsync(out_r, P_.b_r.ptr, n_out * 8, hipMemcpyDeviceToHost, stream()));
    TP_HIP_OK(hipMemcpyAsync(out_i, P_.b_i.ptr, n_out * 8, hipMemcpyDeviceToHost, stream()));
    TP_HIP_OK(hipStreamSynchronize(stream()));
    return true;
}

bool gpu_expand(const u64* in_r, const u64* in_i, size_t n_src, size_t nc, u64* out_r,
                u64* out_i)
{
    Pool& P_ = g_pool;
    const size_t total = n_src * nc;
    if (!h2d(P_.a_r, in_r, n_src * 8) || !h2d(P_.a_i, in_i, n_src * 8)) return false;
    if (!P_.b_r.ensure(total * 8) || !P_.b_i.ensure(total * 8)) return false;
    k_expand<<<grid_for(total), kBlock, 0, stream()>>>(P_.a_r.as<u64>(), P_.a_i.as<u64>(), n_src,
                                                        nc, P_.b_r.as<u64>(), P_.b_i.as<u64>());
    if (!launch_ok()) return false;
    TP_HIP_OK(hipMemcpyAsync(out_r, P_.b_r.ptr, total * 8, hipMemcpyDeviceToHost, stream()));
    TP_HIP_OK(hipMemcpyAsync(out_i, P_.b_i.ptr, total * 8, hipMemcpyDeviceToHost, stream()));
    TP_HIP_OK(hipStreamSynchronize(stream()));
    return true;
}

bool gpu_beta(const u64* z_r, const u64* z_i, int n_z, u64 init_r, u64 init_i, u64* out_r,
              u64* out_i)
{
    Pool& P_ = g_pool;
    const size_t n = (size_t)1 << n_z;
    if (!h2d(P_.z_r, z_r, n_z * 8) || !h2d(P_.z_i, z_i, n_z * 8)) return false;
    if (!P_.b_r.ensure(n * 8) || !P_.b_i.ensure(n * 8)) return false;
    const unsigned blocks = n_z > 8 ? (unsigned)(n >> 8) : 1u;
    k_beta<<<blocks, kBlock, 0, stream()>>>(P_.z_r.as<u64>(), P_.z_i.as<u64>(), n_z, init_r,
                                            init_i, P_.b_r.as<u64>(), P_.b_i.as<u64>());
    if (!launch_ok()) return false;
    TP_HIP_OK(hipMemcpyAsync(out_r, P_.b_r.ptr, n * 8, hipMemcpyDeviceToHost, stream()));
    TP_HIP_OK(hipMemcpyAsync(out_i, P_.b_i.ptr, n * 8, hipMemcpyDeviceToHost, stream()));
    TP_HIP_OK(hipStreamSynchronize(stream()));
    return true;
}

bool gpu_mle(const u64* inputs, const u64* z_r, const u64* z_i, int n_z, u64* out_r, u64* out_i)
{
    Pool& P_ = g_pool;
    const size_t n = (size_t)1 << n_z;
    if (!h2d(P_.mle_in, inputs, n * 8)) return false;
    if (!h2d(P_.z_r, z_r, n_z * 8) || !h2d(P_.z_i, z_i, n_z * 8)) return false;
    const int lb = n_z < 10 ? n_z : 10;
    const size_t rows = n >> lb;
    const unsigned blocks = (unsigned)(rows < 1024 ? rows : 1024);
    RedView rv;
    if (!red_view(P_.red, blocks, 1, rv)) return false;
    k_mle_base<<<blocks, kBlock, 0, stream()>>>(P_.mle_in.as<u64>(), P_.z_r.as<u64>(),
                                                P_.z_i.as<u64>(), n_z, rv.partials);
    if (!launch_ok() || !finalize<1>(rv, blocks)) return false;
    fp2 res;
    TP_HIP_OK(hipMemcpyAsync(&res, rv.result, sizeof(fp2), hipMemcpyDeviceToHost, stream()));
    TP_HIP_OK(hipStreamSynchronize(stream()));
    *out_r = res.re;
    *out_i = res.im;
    return true;
}

bool upload_gates(DevBuf& gt, DevBuf& i0, DevBuf& i1, DevBuf& zr, DevBuf& zi, int n_gates,
                  const int* gate_type, const int* in0, const int* in1, const u64* z1_r,
                  const u64* z1_i)
{
    return h2d(gt, gate_type, n_gates * 4) && h2d(i0, in0, n_gates * 4) &&
           h2d(i1, in1, n_gates * 4) && h2d(zr, z1_r, n_gates * 8) && h2d(zi, z1_i, n_gates * 8);
}

// A session's claim terms (claim_terms.hpp) with their coefficients formed
// here from the caller's z1 (canonical Fp2 sums); returns the term count, or
// -1 with the error recorded.
int upload_terms(DevBuf& gt, DevBuf& i0, DevBuf& i1, DevBuf& zr, DevBuf& zi, int n_gates,
                 const int* gate_type, const int* in0, const int* in1, const u64* z1_r,
                 const u64* z1_i)
{
    const ClaimTerms T = build_claim_terms(n_gates, gate_type, in0, in1);
    const int n = T.n_terms();
    std::vector<u64> tr(n), ti(n);
    for (int k = 0; k < n; ++k) {
        fp2 acc = zero2();
        for (int e = T.off[k]; e < T.off[k + 1]; ++e) {
            const fp2 z = mk2(z1_r[T.src[e]] % P, z1_i[T.src[e]] % P);
            acc = T.neg[e] ? sub2(acc, z) : add2(acc, z);
        }
        tr[k] = acc.re;
        ti[k] = acc.im;
    }
    if (!upload_gates(gt, i0, i1, zr, zi, n, T.type.data(), T.in0.data(), T.in1.data(), tr.data(),
                      ti.data()))
        return -1;
    // the vectors above are host memory the async copies read: wait for them
    if (hipStreamSynchronize(stream()) != hipSuccess) return fail("early_create: upload failed"), -1;
    return n;
}

void store8(const fp2 res[4], u64 out8[8])
{
    for (int t = 0; t < 4; ++t) { out8[2 * t] = res[t].re; out8[2 * t + 1] = res[t].im; }
}

bool gpu_sumcheck_early(int n_gates, const int* gate_type, const int* in0, const int* in1,
                        const u64* z1_r, const u64* z1_i, int n_copies, int n_wires,
                        const u64* v_r, const u64* v_i, const u64* f0_r, const u64* f0_i,
                        const u64* f1_r, const u64* f1_i, const u64* b_r, const u64* b_i,
                        const u64* bf0_r, const u64* bf0_i, const u64* bf1_r, const u64* bf1_i,
                        u64 out8[8])
{
    Pool& P_ = g_pool;
    const size_t vn = (size_t)n_wires * n_copies, hn = (size_t)n_copies / 2;
    const size_t fn = (size_t)n_wires * hn;
    for (int g = 0; g < n_gates; ++g)
        if (in0[g] < 0 || in0[g] >= n_wires || in1[g] < 0 || in1[g] >= n_wires)
            return fail("sumcheck_early: gate %d reads wire outside [0, %d)", g, n_wires);
    if (!upload_gates(P_.gt, P_.i0, P_.i1, P_.z1r, P_.z1i, n_gates, gate_type, in0, in1, z1_r,
                      z1_i))
        return false;
    if (!h2d(P_.v_r, v_r, vn * 8) || !h2d(P_.v_i, v_i, vn * 8) || !h2d(P_.f0_r, f0_r, fn * 8) ||
        !h2d(P_.f0_i, f0_i, fn * 8) || !h2d(P_.f1_r, f1_r, fn * 8) || !h2d(P_.f1_i, f1_i, fn * 8) ||
        !h2d(P_.be_r, b_r, n_copies * 8) || !h2d(P_.be_i, b_i, n_copies * 8) ||
        !h2d(P_.bf0_r, bf0_r, hn * 8) || !h2d(P_.bf0_i, bf0_i, hn * 8) ||
        !h2d(P_.bf1_r, bf1_r, hn * 8) || !h2d(P_.bf1_i, bf1_i, hn * 8))
        return false;
    const int gpt = gates_per_thread(n_gates, (int)hn);
    dim3 grid(grid_for(hn), (unsigned)((n_gates + gpt - 1) / gpt));
    RedView rv;
    if (!red_view(P_.red, (size_t)grid.x * grid.y, 4, rv)) return false;
    k_sumcheck_early<<<grid, kBlock, 0, stream()>>>(
        n_gates, gpt, P_.gt.as<int>(), P_.i0.as<int>(), P_.i1.as<int>(), P_.z1r.as<u64>(),
        P_.z1i.as<u64>(), n_copies, P_.v_r.as<u64>(), P_.v_i.as<u64>(), P_.f0_r.as<u64>(),
        P_.f0_i.as<u64>(), P_.f1_r.as<u64>(), P_.f1_i.as<u64>(), P_.be_r.as<u64>(),
        P_.be_i.as<u64>(), P_.bf0_r.as<u64>(), P_.bf0_i.as<u64>(), P_.bf1_r.as<u64>(),
        P_.bf1_i.as<u64>(), rv.partials);
    if (!launch_ok() || !finalize<4>(rv, (size_t)grid.x * grid.y)) return false;
    fp2 res[4];
    TP_HIP_OK(hipMemcpyAsync(res, rv.result, sizeof(res), hipMemcpyDeviceToHost, stream()));
    TP_HIP_OK(hipStreamSynchronize(stream()));
    store8(res, out8);
    return true;
}

// ------------------------------------------------------- early session --

struct EarlySession {
    int n_gates = 0, n_wires = 0, n_copies = 0;
    int n_terms = 0;                      // claim terms (claim_terms.hpp) in gt, i0, i1, z1r, z1i
    bool ping = false;                    // which of (val, val2) is current
    DevBuf gt, i0, i1, z1r, z1i;
    DevBuf val_r, val_i, val2_r, val2_i;
    DevBuf beta_r, beta_i, beta2_r, beta2_i;
    DevBuf red;
    ~EarlySession()
    {
        DevBuf* all[] = {&gt, &i0, &i1, &z1r, &z1i, &val_r, &val_i, &val2_r, &val2_i,
                         &beta_r, &beta_i, &beta2_r, &beta2_i, &red};
        for (DevBuf* b : all) b->release();
    }
    u64* cur_vr() { return (ping ? val2_r : val_r).as<u64>(); }
    u64* cur_vi() { return (ping ? val2_i : val_i).as<u64>(); }
    u64* cur_br() { return (ping ? beta2_r : beta_r).as<u64>(); }
    u64* cur_bi() { return (ping ? beta2_i : beta_i).as<u64>(); }
    u64* nxt_vr() { return (ping ? val_r : val2_r).as<u64>(); }
    u64* nxt_vi() { return (ping ? val_i : val2_i).as<u64>(); }
    u64* nxt_br() { return (ping ? beta_r : beta2_r).as<u64>(); }
    u64* nxt_bi() { return (ping ? beta_i : beta2_i).as<u64>(); }
};

bool session_claim(EarlySession* s, u64 out8[8])
{
    const int n_pairs = s->n_copies >> 1;
    if (n_pairs < 1) return fail("early_claim: session already folded to one copy");
    const ClaimShape cs = claim_shape(s->n_terms, n_pairs, kClaimItems);
    RedView rv;
    if (!red_view(s->red, cs.blocks, 4, rv)) return false;
    kern::k_early_claim<<<cs.blocks, kBlock, 0, stream()>>>(
        s->n_terms, cs.gpt, cs.n_chunks, s->gt.as<int>(), s->i0.as<int>(), s->i1.as<int>(),
        s->z1r.as<u64>(), s->z1i.as<u64>(), s->n_copies, s->cur_vr(), s->cur_vi(), s->cur_br(),
        s->cur_bi(), rv.partials);
    if (!launch_ok() || !finalize<4>(rv, cs.blocks)) return false;
    fp2 res[4];
    TP_HIP_OK(hipMemcpyAsync(res, rv.result, sizeof(res), hipMemcpyDeviceToHost, stream()));
    TP_HIP_OK(hipStreamSynchronize(stream()));
    store8(res, out8);
    return true;
}

bool session_fold(EarlySession* s, u64 rr, u64 ri)
{
    const int n_out = s->n_copies >> 1;
    if (n_out < 1) return fail("early_fold: session already folded to one copy");
    int log_out = 0;
    while ((1 << log_out) < n_out) ++log_out;
    const size_t total = ((size_t)s->n_wires + 1) * n_out;
    kern::k_fold_rows<<<grid_for(total), kBlock, 0, stream()>>>(
        s->cur_vr(), s->cur_vi(), s->n_wires, log_out, rr, ri, s->nxt_vr(), s->nxt_vi(),
        s->cur_br(), s->cur_bi(), s->nxt_br(), s->nxt_bi());
    if (!launch_ok()) return false;
    TP_HIP_OK(hipStreamSynchronize(stream()));
    s->ping = !s->ping;
    s->n_copies = n_out;
    return true;
}

bool session_finish(EarlySession* s, u64* w_r, u64* w_i, u64* b_r, u64* b_i)
{
    const size_t stride = (size_t)s->n_copies * 8;
    // element 0 of every wire row: one strided 2-D copy per coordinate
    TP_HIP_OK(hipMemcpy2DAsync(w_r, 8, s->cur_vr(), stride, 8, s->n_wires,
                               hipMemcpyDeviceToHost, stream()));
    TP_HIP_OK(hipMemcpy2DAsync(w_i, 8, s->cur_vi(), stride, 8, s->n_wires,
                               hipMemcpyDeviceToHost, stream()));
    TP_HIP_OK(hipMemcpyAsync(b_r, s->cur_br(), 8, hipMemcpyDeviceToHost, stream()));
    TP_HIP_OK(hipMemcpyAsync(b_i, s->cur_bi(), 8, hipMemcpyDeviceToHost, stream()));
    TP_HIP_OK(hipStreamSynchronize(stream()));
    return true;
}

EarlySession* session_create(int n_gates, const int* gate_type, const int* in0, const int* in1,
                             const u64* z1_r, const u64* z1_i, int n_copies, int n_wires,
                             const u64* v_r, const u64* v_i, const u64* b_r, const u64* b_i)
{
    // Same rejections as TP/native_fp2/native_fp2_core.hpp:181-182.
    if (n_gates <= 0 || n_wires <= 0 || n_copies < 2 || (n_copies & (n_copies - 1)) != 0) {
        fail("early_create: bad shape (n_gates=%d n_wires=%d n_copies=%d)", n_gates, n_wires,
             n_copies);
        return nullptr;
    }
    for (int g = 0; g < n_gates; ++g) {
        if (in0[g] < 0 || in0[g] >= n_wires || in1[g] < 0 || in1[g] >= n_wires) {
            fail("early_create: gate %d reads wire outside [0, %d)", g, n_wires);
            return nullptr;
        }
    }
    auto* s = new EarlySession();
    s->n_gates = n_gates;
    s->n_wires = n_wires;
    s->n_copies = n_copies;
    const size_t vn = (size_t)n_wires * n_copies;
    s->n_terms = upload_terms(s->gt, s->i0, s->i1, s->z1r, s->z1i, n_gates, gate_type, in0, in1,
                              z1_r, z1_i);
    bool ok = s->n_terms > 0 &&
              h2d(s->val_r, v_r, vn * 8) && h2d(s->val_i, v_i, vn * 8) &&
              s->val2_r.ensure(vn * 8 / 2) && s->val2_i.ensure(vn * 8 / 2) &&
              h2d(s->beta_r, b_r, n_copies * 8) && h2d(s->beta_i, b_i, n_copies * 8) &&
              s->beta2_r.ensure(n_copies * 8 / 2) && s->beta2_i.ensure(n_copies * 8 / 2);
    if (ok && hipStreamSynchronize(stream()) != hipSuccess) ok = fail("early_create: upload failed");
    if (!ok) {
        delete s;
        return nullptr;
    }
    return s;
}

// Session over wire tables already on the device (SoA [wire][copy]): the
// tables are copied device-to-device, beta and the gate arrays come from the host.
EarlySession* session_create_dev(int n_gates, const int* gate_type, const int* in0,
                                 const int* in1, const u64* z1_r, const u64* z1_i, int n_copies,
                                 int n_wires, const u64* d_v_r, const u64* d_v_i, const u64* b_r,
                                 const u64* b_i)
{
    if (n_gates <= 0 || n_wires <= 0 || n_copies < 2 || (n_copies & (n_copies - 1)) != 0) {
        fail("early_create_dev: bad shape (n_gates=%d n_wires=%d n_copies=%d)", n_gates, n_wires,
             n_copies);
        return nullptr;
    }
    for (int g = 0; g < n_gates; ++g)
        if (in0[g] < 0 || in0[g] >= n_wires || in1[g] < 0 || in1[g] >= n_wires) {
            fail("early_create_dev: gate %d reads wire outside [0, %d)", g, n_wires);
            return nullptr;
        }
    auto* s = new EarlySession();
    s->n_gates = n_gates;
    s->n_wires = n_wires;
    s->n_copies = n_copies;
    const size_t vn = (size_t)n_wires * n_copies;
    s->n_terms = upload_terms(s->gt, s->i0, s->i1, s->z1r, s->z1i, n_gates, gate_type, in0, in1,
                              z1_r, z1_i);
    bool ok = s->n_terms > 0 &&
              s->val_r.ensure(vn * 8) && s->val_i.ensure(vn * 8) &&
              s->val2_r.ensure(vn * 8 / 2) && s->val2_i.ensure(vn * 8 / 2) &&
              h2d(s->beta_r, b_r, n_copies * 8) && h2d(s->beta_i, b_i, n_copies * 8) &&
              s->beta2_r.ensure(n_copies * 8 / 2) && s->beta2_i.ensure(n_copies * 8 / 2);
    ok = ok && hipMemcpyAsync(s->val_r.ptr, d_v_r, vn * 8, hipMemcpyDeviceToDevice, stream()) == hipSuccess &&
         hipMemcpyAsync(s->val_i.ptr, d_v_i, vn * 8, hipMemcpyDeviceToDevice, stream()) == hipSuccess;
    if (ok && hipStreamSynchronize(stream()) != hipSuccess) ok = fail("early_create_dev: copy failed");
    if (!ok) {
        if (!tp_last_error_set()) fail("early_create_dev: device copy failed");
        delete s;
        return nullptr;
    }
    return s;
}

}  // namespace

// ===================================================================== ABI ==

extern "C" {

int native_fp2_init(void) { return tp::device_ready() ? 0 : -1; }

int native_fp2_cuda_available(void) { return tp::device_ready() ? 1 : 0; }

void native_fp2_fold(const uint64_t* in_real, const uint64_t* in_imag, size_t n_out,
                     uint64_t r_real, uint64_t r_imag, uint64_t* out_real, uint64_t* out_imag)
{
    tp::clear_error();
    if (n_out == 0 || !tp::device_ready()) return;
    gpu_fold(in_real, in_imag, n_out, r_real, r_imag, out_real, out_imag);
}

void native_fp2_expand(const uint64_t* in_real, const uint64_t* in_imag, size_t n_src,
                       size_t ncopies, uint64_t* out_real, uint64_t* out_imag)
{
    tp::clear_error();
    if (n_src == 0 || ncopies == 0 || !tp::device_ready()) return;
    gpu_expand(in_real, in_imag, n_src, ncopies, out_real, out_imag);
}

void native_fp2_sumcheck_early(int n_gates, const int* gate_type, const int* in0, const int* in1,
                               const uint64_t* z1_real, const uint64_t* z1_imag, int n_copies,
                               int n_wires, const uint64_t* values_real,
                               const uint64_t* values_imag, const uint64_t* fact0_real,
                               const uint64_t* fact0_imag, const uint64_t* fact1_real,
                               const uint64_t* fact1_imag, const uint64_t* beta_real,
                               const uint64_t* beta_imag, const uint64_t* beta_f0_real,
                               const uint64_t* beta_f0_imag, const uint64_t* beta_f1_real,
                               const uint64_t* beta_f1_imag, uint64_t out8[8])
{
    tp::clear_error();
    if (n_gates <= 0 || n_copies < 2) {
        memset(out8, 0, 8 * sizeof(uint64_t));   // empty sums, as the reference loop yields
        return;
    }
    if (!tp::device_ready()) return;
    gpu_sumcheck_early(n_gates, gate_type, in0, in1, z1_real, z1_imag, n_copies, n_wires,
                       values_real, values_imag, fact0_real, fact0_imag, fact1_real, fact1_imag,
                       beta_real, beta_imag, beta_f0_real, beta_f0_imag, beta_f1_real,
                       beta_f1_imag, out8);
}

void native_fp2_compute_beta(const uint64_t* z_real, const uint64_t* z_imag, int n_z,
                             uint64_t init_real, uint64_t init_imag, uint64_t* out_real,
                             uint64_t* out_imag)
{
    tp::clear_error();
    if (n_z <= 0) return;   // reference writes nothing for n_z <= 0 (core :144)
    if (n_z > 40) { tp::fail("compute_beta: n_z=%d too large", n_z); return; }
    if (!tp::device_ready()) return;
    gpu_beta(z_real, z_imag, n_z, init_real, init_imag, out_real, out_imag);
}

void native_fp2_mle_eval_base(const uint64_t* inputs, const uint64_t* z_real,
                              const uint64_t* z_imag, int n_z, uint64_t* out_real,
                              uint64_t* out_imag)
{
    tp::clear_error();
    if (n_z <= 0) {   // reference returns inputs[0] as stored (core :304-307)
        *out_real = inputs ? inputs[0] : 0;
        *out_imag = 0;
        return;
    }
    if (n_z > 40) { tp::fail("mle_eval_base: n_z=%d too large", n_z); return; }
    if (!tp::device_ready()) return;
    gpu_mle(inputs, z_real, z_imag, n_z, out_real, out_imag);
}

void* native_fp2_early_create(int n_gates, const int* gate_type, const int* in0, const int* in1,
                              const uint64_t* z1_real, const uint64_t* z1_imag, int n_copies,
                              int n_wires, const uint64_t* values_real,
                              const uint64_t* values_imag, const uint64_t* beta_real,
                              const uint64_t* beta_imag)
{
    tp::clear_error();
    if (!tp::device_ready()) return nullptr;
    return session_create(n_gates, gate_type, in0, in1, z1_real, z1_imag, n_copies, n_wires,
                          values_real, values_imag, beta_real, beta_imag);
}

int native_fp2_early_claim(void* handle, uint64_t out8[8])
{
    tp::clear_error();
    auto* s = reinterpret_cast<EarlySession*>(handle);
    if (!s) return -1;
    return session_claim(s, out8) ? 0 : -1;
}

int native_fp2_early_fold(void* handle, uint64_t r_real, uint64_t r_imag)
{
    tp::clear_error();
    auto* s = reinterpret_cast<EarlySession*>(handle);
    if (!s) return -1;
    return session_fold(s, r_real, r_imag) ? 0 : -1;
}

int native_fp2_early_finish(void* handle, uint64_t* wire_real, uint64_t* wire_imag,
                            uint64_t* beta_real, uint64_t* beta_imag)
{
    tp::clear_error();
    auto* s = reinterpret_cast<EarlySession*>(handle);
    if (!s) return -1;
    return session_finish(s, wire_real, wire_imag, beta_real, beta_imag) ? 0 : -1;
}

int native_fp2_early_n_copies(void* handle)
{
    auto* s = reinterpret_cast<EarlySession*>(handle);
    return s ? s->n_copies : 0;
}

int native_fp2_early_n_wires(void* handle)
{
    auto* s = reinterpret_cast<EarlySession*>(handle);
    return s ? s->n_wires : 0;
}

void native_fp2_early_destroy(void* handle) { delete reinterpret_cast<EarlySession*>(handle); }

// Early session over device-resident SoA wire tables [n_wires][n_copies]
// (e.g. from tp_wire_table); otherwise as native_fp2_early_create.
void* tp_early_create_dev(int n_gates, const int* gate_type, const int* in0, const int* in1,
                          const uint64_t* z1_real, const uint64_t* z1_imag, int n_copies,
                          int n_wires, const uint64_t* d_values_real, const uint64_t* d_values_imag,
                          const uint64_t* beta_real, const uint64_t* beta_imag)
{
    tp::clear_error();
    if (!tp::device_ready()) return nullptr;
    return session_create_dev(n_gates, gate_type, in0, in1, z1_real, z1_imag, n_copies, n_wires,
                              d_values_real, d_values_imag, beta_real, beta_imag);
}

}  // extern "C"
