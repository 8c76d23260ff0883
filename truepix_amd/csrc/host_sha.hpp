// host_sha.hpp - SHA-256 (FIPS 180-4) for host code, with the x86 SHA
// extensions when the CPU has them (checked once at run time) and a portable
// compression function otherwise.  The frame prover absorbs megabytes of
// circuit outputs into the Fiat-Shamir transcript per frame
// (TP/libTruePix/fiatshamir.py:105-114 via circuitnizkvec_vole.py:286-300), so
// the transcript hash runs at the SHA-NI rate (about 2 GB/s per core) instead
// of a scalar one.
//
// Same streaming interface as tp::Sha256 (init / update / peek): peek() digests
// a copy and leaves the running state open, which is how the transcript draws
// challenges from a prefix of its byte stream.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace tp {

// Compresses n 64-byte blocks into st (eight big-endian words, FIPS order).
void sha256_blocks(uint32_t st[8], const uint8_t* data, size_t n);
// True when sha256_blocks uses the SHA extensions.
bool sha256_has_ni();

struct HostSha256 {
    uint32_t h[8];
    uint8_t buf[64];
    uint32_t nbuf;
    uint64_t total;

    void init();
    void update(const uint8_t* p, size_t n);
    void peek(uint8_t out[32]) const;
};

}  // namespace tp
