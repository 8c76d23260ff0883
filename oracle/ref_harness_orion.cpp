// Test-infrastructure harness: the TruePix signer's Orion commit of a frame's
// inputs (OZ/examples/multivariate_PC_test_commit.cpp:265-331, calling
// commit_pair, OZ/src/VPD/linearPC.cpp:253-459), computed with the
// reference's own pieces, compiled in place from /root/reference by
// oracle/Makefile:
//   - the layout and the seeded mask of OZ/include/truepix/truepix_gkr.h
//     (make_layout, alloc_masked_from_seed, alloc_mask_from_seed, root_hex);
//   - expander_init and encode of OZ/include/linear_code/expanders.h and
//     linear_code_encode.h;
//   - the row hash and create_tree of OZ/src/infrastructure/merkle_tree.cpp
//     over flo-shani's SHA-256.
// commit_pair itself lives in linearPC.cpp, which pulls in the whole Orion
// prover; its per-column flow (encode every column, absorb each column pair
// into every row's Merkle-Damgard state, build the tree over the rows) is
// followed here through those functions.  Never part of the product.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <unistd.h>
#include <vector>

#include "infrastructure/merkle_tree.h"
#include "linear_code/linear_code_encode.h"
#include "truepix/orion_fs.h"
#include "truepix/orion_proof.h"
#include "truepix/truepix_gkr.h"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iterator>
#include <string>

extern "C" {

// inputs: [copies][384] sample values (the lines of input.txt).  Writes the two
// roots as root_hex strings (64 hex digits + NUL) and the codeword length of a
// column.  0 / -1.
int ref_orion_commit(int copies, const uint8_t* inputs, uint64_t seed, char masked_root[65],
                     char mask_root[65], long long* cw_len_out)
{
    truepix::Layout lay;
    if (!truepix::make_layout(copies, lay)) return -1;
    std::vector<prime_field::field_element> coefs(lay.n, prime_field::field_element(0ULL));
    for (int c = 0; c < copies; ++c) {
        const long long base = (long long)c * lay.padded_copy_size;
        for (int w = 0; w < lay.copy_size; ++w)
            coefs[base + w] = prime_field::field_element((unsigned long long)inputs[(long long)c * lay.copy_size + w]);
        if (lay.const_idx >= 0)
            coefs[base + lay.const_idx] = prime_field::field_element((unsigned long long)lay.const_val);
    }
    // the signer process reaches expander_init on glibc's initial rand() state;
    // encode's scratch is sized on its first call, so it is re-sized per call
    srand(1);
    __encode_initialized = false;
    expander_init(lay.n / column_size);
    prime_field::field_element* polys[2] = {truepix::alloc_masked_from_seed(coefs.data(), lay, seed),
                                            truepix::alloc_mask_from_seed(lay, seed)};
    const long long col_len = lay.n / column_size;
    const long long rows = col_len * 2;
    const long long cw_len = encode_output_len(col_len);
    char* roots[2] = {masked_root, mask_root};
    for (int k = 0; k < 2; ++k) {
        std::vector<__hhash_digest> stash(rows);
        memset(stash.data(), 0, sizeof(__hhash_digest) * (size_t)rows);
        std::vector<prime_field::field_element> c0(rows), c1(rows);
        for (int j = 0; j < column_size / 2; ++j) {
            for (long long r = 0; r < rows; ++r) c0[r] = c1[r] = prime_field::field_element(0ULL);
            encode(&polys[k][(long long)(2 * j) * col_len], c0.data(), col_len);
            encode(&polys[k][(long long)(2 * j + 1) * col_len], c1.data(), col_len);
            for (long long r = 0; r < rows; ++r)
                stash[r] = merkle_tree::hash_double_field_element_merkle_damgard(c0[r], c1[r], stash[r]);
        }
        __hhash_digest* mt = nullptr;
        merkle_tree::merkle_tree_prover::create_tree(stash.data(), (int)rows, mt,
                                                    sizeof(__hhash_digest), true);
        const std::string hex = truepix::root_hex(mt[1]);
        memcpy(roots[k], hex.c_str(), 64);
        roots[k][64] = 0;
        free(mt);
    }
    delete[] polys[0];
    delete[] polys[1];
    if (cw_len_out) *cw_len_out = cw_len;
    return 0;
}

// One column's codeword (encode of the reference), for checking the encoder
// alone: src n Fp2 as (real, imag) pairs, dst (real, imag) pairs; the
// expander graphs are regenerated from glibc's initial state for size n.
long long ref_orion_encode(long long n, const uint64_t* src, uint64_t* dst)
{
    srand(1);
    __encode_initialized = false;
    expander_init(n);
    std::vector<prime_field::field_element> s(n), d(2 * n + 64);
    for (long long i = 0; i < n; ++i) {
        s[i].real = src[2 * i];
        s[i].img = src[2 * i + 1];
    }
    const int len = encode(s.data(), d.data(), n);
    for (int i = 0; i < len; ++i) {
        dst[2 * i] = d[i].real;
        dst[2 * i + 1] = d[i].img;
    }
    return len;
}

// The opening of one committed polynomial, computed as tensor_product_protocol
// (linearPC.cpp:700-966) computes its OpeningProof, from the reference's
// encode, row hash, create_tree, hash_single_field_element and
// collect_merkle_siblings, serialized by OpeningProof::write_file into a
// temporary file and read back.  r0 / r1 are formed as open_and_verify's dfs
// does (linearPC.cpp:996-1034; dfs is a static helper of linearPC.cpp, which
// pulls in the whole prover, so its three lines are followed here).  The
// proof bytes go to a malloc'd buffer (free with ref_free).  0 / -1.
static void ref_dfs(prime_field::field_element* dst, const prime_field::field_element* r, long long size,
                    int depth, prime_field::field_element val)
{
    if (size == 1) {
        *dst = val;
        return;
    }
    ref_dfs(dst, r, size / 2, depth + 1, val * (prime_field::field_element(1) - r[depth]));
    ref_dfs(dst + size / 2, r, size / 2, depth + 1, val * r[depth]);
}

int ref_orion_open(int copies, const uint8_t* inputs, uint64_t seed, int use_group1,
                   const uint64_t* r, int n_r, const long long* col_q, int nq, uint8_t** proof,
                   size_t* proof_len)
{
    truepix::Layout lay;
    if (!truepix::make_layout(copies, lay)) return -1;
    if (n_r != lay.lg_n) return -1;
    std::vector<prime_field::field_element> coefs(lay.n, prime_field::field_element(0ULL));
    for (int c = 0; c < copies; ++c) {
        const long long base = (long long)c * lay.padded_copy_size;
        for (int w = 0; w < lay.copy_size; ++w)
            coefs[base + w] = prime_field::field_element((unsigned long long)inputs[(long long)c * lay.copy_size + w]);
        if (lay.const_idx >= 0)
            coefs[base + lay.const_idx] = prime_field::field_element((unsigned long long)lay.const_val);
    }
    srand(1);
    __encode_initialized = false;
    expander_init(lay.n / column_size);
    prime_field::field_element* poly = use_group1 ? truepix::alloc_masked_from_seed(coefs.data(), lay, seed)
                                                  : truepix::alloc_mask_from_seed(lay, seed);
    const long long msg_len = lay.n / column_size, rows = msg_len * 2;
    const long long cw_len = encode_output_len(msg_len);
    std::vector<std::vector<prime_field::field_element>> cws(column_size);
    for (int j = 0; j < column_size; ++j) {
        cws[j].assign(rows, prime_field::field_element(0ULL));
        encode(&poly[(long long)j * msg_len], cws[j].data(), msg_len);
    }
    std::vector<__hhash_digest> stash(rows);
    memset(stash.data(), 0, sizeof(__hhash_digest) * (size_t)rows);
    for (int j = 0; j < column_size / 2; ++j)
        for (long long rr = 0; rr < rows; ++rr)
            stash[rr] = merkle_tree::hash_double_field_element_merkle_damgard(cws[2 * j][rr], cws[2 * j + 1][rr],
                                                                              stash[rr]);
    __hhash_digest* com_mt = nullptr;
    merkle_tree::merkle_tree_prover::create_tree(stash.data(), (int)rows, com_mt, sizeof(__hhash_digest), true);
    std::vector<prime_field::field_element> rv(n_r);
    for (int d = 0; d < n_r; ++d) {
        rv[d].real = r[2 * d];
        rv[d].img = r[2 * d + 1];
    }
    int lc = 0;
    while ((1 << lc) < column_size) ++lc;
    std::vector<prime_field::field_element> r0(column_size), r1(msg_len);
    ref_dfs(r0.data(), rv.data(), column_size, 0, prime_field::field_element(1));
    ref_dfs(r1.data(), rv.data() + lc, msg_len, 0, prime_field::field_element(1));
    std::vector<prime_field::field_element> cc(cw_len, prime_field::field_element(0ULL)),
        cm(msg_len, prime_field::field_element(0ULL)), qc((size_t)nq * column_size);
    for (int j = 0; j < column_size; ++j) {
        for (long long t = 0; t < cw_len; ++t) cc[t] = cc[t] + r0[j] * cws[j][t];
        for (int i = 0; i < nq; ++i) qc[(size_t)i * column_size + j] = cws[j][col_q[i]];
        for (long long t = 0; t < msg_len; ++t) cm[t] = cm[t] + r0[j] * poly[(long long)j * msg_len + t];
    }
    prime_field::field_element answer(0ULL);
    for (long long t = 0; t < msg_len; ++t) answer = answer + r1[t] * cm[t];
    std::vector<__hhash_digest> leaves(rows);
    prime_field::field_element zero;
    memset(&zero, 0, sizeof(zero));
    for (long long i = 0; i < rows; ++i)
        leaves[i] = merkle_tree::hash_single_field_element(i < cw_len ? cc[i] : zero);
    __hhash_digest* cmb_mt = nullptr;
    merkle_tree::merkle_tree_prover::create_tree(leaves.data(), (int)rows, cmb_mt, sizeof(__hhash_digest), true);
    truepix_orion::OpeningProof op;
    op.version = 2;
    op.use_group1 = use_group1 ? 1u : 0u;
    op.N = (uint64_t)lay.n;
    op.column_size = (uint64_t)column_size;
    op.msg_len = (uint64_t)msg_len;
    op.cw_len = (uint64_t)cw_len;
    op.query_count = (uint32_t)nq;
    op.answer = answer;
    op.com_root = com_mt[1];
    op.combined_root = cmb_mt[1];
    op.col_q.assign(col_q, col_q + nq);
    op.queried_cols = qc;
    op.combined_codeword = cc;
    op.combined_message = cm;
    truepix_orion::collect_merkle_siblings(com_mt, (int)rows, col_q, nq, op.com_mt_idx, op.com_mt);
    truepix_orion::collect_merkle_siblings(cmb_mt, (int)rows, col_q, nq, op.combined_mt_idx, op.combined_mt);
    char path[] = "/tmp/ref_orion_open_XXXXXX";
    const int fd = mkstemp(path);
    if (fd < 0) return -1;
    close(fd);
    const bool ok = op.write_file(path);
    std::ifstream is(path, std::ios::binary);
    std::vector<char> bytes((std::istreambuf_iterator<char>(is)), std::istreambuf_iterator<char>());
    std::remove(path);
    free(com_mt);
    free(cmb_mt);
    delete[] poly;
    if (!ok) return -1;
    *proof = static_cast<uint8_t*>(malloc(bytes.size()));
    memcpy(*proof, bytes.data(), bytes.size());
    *proof_len = bytes.size();
    return 0;
}

// The queries tensor_product_protocol draws in NIZK mode (linearPC.cpp:711-758)
// from the reference's own transcript (orion_fs.cpp's global_fs), for the
// opened root (digest bytes), N, the group and the point r; q gets
// query_count positions.  Returns query_count.
int ref_orion_draw_queries(const uint8_t root[32], long long n, int use_group1, long long cw_len,
                           const uint64_t* r, int n_r, long long* q)
{
    std::vector<prime_field::field_element> rv(n_r);
    for (int d = 0; d < n_r; ++d) {
        rv[d].real = r[2 * d];
        rv[d].img = r[2 * d + 1];
    }
    int lc = 0;
    while ((1 << lc) < column_size) ++lc;
    const long long msg_len = n / column_size;
    std::vector<prime_field::field_element> r0(column_size), r1(msg_len);
    ref_dfs(r0.data(), rv.data(), column_size, 0, prime_field::field_element(1));
    ref_dfs(r1.data(), rv.data() + lc, msg_len, 0, prime_field::field_element(1));
    __hhash_digest com;
    memcpy(&com, root, 32);
    truepix_orion::fs_reset_global();
    auto& fs = truepix_orion::global_fs();
    fs.absorb_label("tensor-open");
    fs.absorb_digest(com);
    fs.absorb_u64(static_cast<uint64_t>(n));
    fs.absorb_u64(use_group1 ? 1ULL : 0ULL);
    for (long long i = 0; i < column_size; ++i) fs.absorb_field(r0[i]);
    for (long long i = 0; i < msg_len; ++i) fs.absorb_field(r1[i]);
    const int query_count = -128 / (log2(1 - target_distance));
    std::vector<int> group;
    for (long long i = 0; i < cw_len; ++i)
        if ((i % 2 == 0) == (use_group1 != 0)) group.push_back((int)i);
    for (int i = 0; i < query_count; ++i)
        q[i] = group[truepix_orion::orion_challenge_index((int)group.size())];
    return query_count;
}

void ref_free(void* p) { free(p); }

}  // extern "C"
