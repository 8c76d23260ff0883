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
#include <vector>

#include "infrastructure/merkle_tree.h"
#include "linear_code/linear_code_encode.h"
#include "truepix/truepix_gkr.h"

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

}  // extern "C"
