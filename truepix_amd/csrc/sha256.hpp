// sha256.hpp - SHA-256 (FIPS 180-4), usable from host code and device code.
// Streaming state: absorb bytes, then digest a copy (the running state is not
// finalized), which is how the TruePix transcript draws challenges from a
// prefix of its byte stream (TP/libTruePix/fiatshamir.py:144-152 calls
// hexdigest() on the live hash object).
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifndef TP_HD
#include <hip/hip_runtime.h>
#define TP_HD __host__ __device__ __forceinline__
#endif

namespace tp {

struct Sha256 {
    uint32_t h[8];
    uint8_t buf[64];
    uint32_t nbuf;
    uint64_t total;   // bytes absorbed

    TP_HD static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

    TP_HD void init()
    {
        const uint32_t iv[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                                0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
        for (int i = 0; i < 8; ++i) h[i] = iv[i];
        nbuf = 0;
        total = 0;
    }

    TP_HD static void compress(uint32_t* st, const uint8_t* blk)
    {
        const uint32_t K[64] = {
            0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u,
            0x923f82a4u, 0xab1c5ed5u, 0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u,
            0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u,
            0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
            0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u,
            0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u,
            0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u, 0xa2bfe8a1u, 0xa81a664bu,
            0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
            0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au,
            0x5b9cca4fu, 0x682e6ff3u, 0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u,
            0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};
        uint32_t w[64];
        for (int i = 0; i < 16; ++i)
            w[i] = ((uint32_t)blk[4 * i] << 24) | ((uint32_t)blk[4 * i + 1] << 16) |
                   ((uint32_t)blk[4 * i + 2] << 8) | (uint32_t)blk[4 * i + 3];
        for (int i = 16; i < 64; ++i) {
            const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
            const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        uint32_t a = st[0], b = st[1], c = st[2], d = st[3], e = st[4], f = st[5], g = st[6],
                 hh = st[7];
        for (int i = 0; i < 64; ++i) {
            const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
            const uint32_t ch = (e & f) ^ (~e & g);
            const uint32_t t1 = hh + S1 + ch + K[i] + w[i];
            const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
            const uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
            const uint32_t t2 = S0 + mj;
            hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
        }
        st[0] += a; st[1] += b; st[2] += c; st[3] += d;
        st[4] += e; st[5] += f; st[6] += g; st[7] += hh;
    }

    TP_HD void update(const uint8_t* p, size_t n)
    {
        total += n;
        while (n) {
            uint32_t take = 64 - nbuf;
            if (take > n) take = (uint32_t)n;
            for (uint32_t i = 0; i < take; ++i) buf[nbuf + i] = p[i];
            nbuf += take;
            p += take;
            n -= take;
            if (nbuf == 64) { compress(h, buf); nbuf = 0; }
        }
    }

    // Digest of everything absorbed so far; the running state is unchanged.
    TP_HD void peek(uint8_t out[32]) const
    {
        uint32_t st[8];
        uint8_t blk[64];
        for (int i = 0; i < 8; ++i) st[i] = h[i];
        for (uint32_t i = 0; i < nbuf; ++i) blk[i] = buf[i];
        uint32_t m = nbuf;
        blk[m++] = 0x80;
        if (m > 56) {
            while (m < 64) blk[m++] = 0;
            compress(st, blk);
            m = 0;
        }
        while (m < 56) blk[m++] = 0;
        const uint64_t bits = total * 8;
        for (int i = 0; i < 8; ++i) blk[56 + i] = (uint8_t)(bits >> (56 - 8 * i));
        compress(st, blk);
        for (int i = 0; i < 8; ++i) {
            out[4 * i] = (uint8_t)(st[i] >> 24);
            out[4 * i + 1] = (uint8_t)(st[i] >> 16);
            out[4 * i + 2] = (uint8_t)(st[i] >> 8);
            out[4 * i + 3] = (uint8_t)st[i];
        }
    }
};

}  // namespace tp
