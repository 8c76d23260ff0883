// host_sha.cpp - SHA-256 for host code (FIPS 180-4), SHA-NI path plus a
// portable fallback.  Host-only; compiled with the system C++ compiler.
#include "host_sha.hpp"

#include <cpuid.h>
#include <immintrin.h>
#include <string.h>

namespace tp {

namespace {

const uint32_t kK[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u,
    0xab1c5ed5u, 0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu,
    0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu,
    0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau, 0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u,
    0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu,
    0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u, 0xa2bfe8a1u, 0xa81a664bu,
    0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u, 0x19a4c116u,
    0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u,
    0xc67178f2u};

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

void blocks_portable(uint32_t st[8], const uint8_t* p, size_t n)
{
    for (; n; --n, p += 64) {
        uint32_t w[64];
        for (int t = 0; t < 16; ++t)
            w[t] = (uint32_t)p[4 * t] << 24 | (uint32_t)p[4 * t + 1] << 16 |
                   (uint32_t)p[4 * t + 2] << 8 | (uint32_t)p[4 * t + 3];
        for (int t = 16; t < 64; ++t) {
            const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
            const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
            w[t] = w[t - 16] + s0 + w[t - 7] + s1;
        }
        uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
        uint32_t e = st[4], f = st[5], g = st[6], h = st[7];
        for (int t = 0; t < 64; ++t) {
            const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
            const uint32_t ch = (e & f) ^ (~e & g);
            const uint32_t t1 = h + S1 + ch + kK[t] + w[t];
            const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
            const uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
            const uint32_t t2 = S0 + mj;
            h = g; g = f; f = e; e = d + t1;
            d = c; c = b; b = a; a = t1 + t2;
        }
        st[0] += a; st[1] += b; st[2] += c; st[3] += d;
        st[4] += e; st[5] += f; st[6] += g; st[7] += h;
    }
}

// SHA extensions: the state lives as (A,B,E,F) and (C,D,G,H) lane groups;
// each sha256rnds2 does two rounds, and the message schedule for 4-word group
// i >= 4 is msg2(msg1(W[i-4], W[i-3]) + (W[i-2]:W[i-1] shifted by one word),
// W[i-1]), which is w[t] = s1(w[t-2]) + w[t-7] + s0(w[t-15]) + w[t-16].
__attribute__((target("sha,sse4.1,ssse3")))
void blocks_ni(uint32_t st[8], const uint8_t* p, size_t n)
{
    const __m128i bswap = _mm_set_epi64x(0x0c0d0e0f08090a0bULL, 0x0405060700010203ULL);
    __m128i tmp = _mm_loadu_si128((const __m128i*)&st[0]);    // A B C D
    __m128i s1 = _mm_loadu_si128((const __m128i*)&st[4]);     // E F G H
    tmp = _mm_shuffle_epi32(tmp, 0xB1);                       // B A D C
    s1 = _mm_shuffle_epi32(s1, 0x1B);                         // H G F E
    __m128i s0 = _mm_alignr_epi8(tmp, s1, 8);                 // F E B A
    s1 = _mm_blend_epi16(s1, tmp, 0xF0);                      // H G D C
    for (; n; --n, p += 64) {
        const __m128i save0 = s0, save1 = s1;
        __m128i w[16];
        for (int i = 0; i < 4; ++i)
            w[i] = _mm_shuffle_epi8(_mm_loadu_si128((const __m128i*)(p + 16 * i)), bswap);
        for (int i = 4; i < 16; ++i)
            w[i] = _mm_sha256msg2_epu32(
                _mm_add_epi32(_mm_sha256msg1_epu32(w[i - 4], w[i - 3]),
                              _mm_alignr_epi8(w[i - 1], w[i - 2], 4)),
                w[i - 1]);
        for (int i = 0; i < 16; ++i) {
            __m128i m = _mm_add_epi32(w[i], _mm_loadu_si128((const __m128i*)&kK[4 * i]));
            s1 = _mm_sha256rnds2_epu32(s1, s0, m);
            m = _mm_shuffle_epi32(m, 0x0E);
            s0 = _mm_sha256rnds2_epu32(s0, s1, m);
        }
        s0 = _mm_add_epi32(s0, save0);
        s1 = _mm_add_epi32(s1, save1);
    }
    tmp = _mm_shuffle_epi32(s0, 0x1B);                        // A B E F
    s1 = _mm_shuffle_epi32(s1, 0xB1);                         // G H C D
    s0 = _mm_blend_epi16(tmp, s1, 0xF0);                      // A B C D
    s1 = _mm_alignr_epi8(s1, tmp, 8);                         // E F G H
    _mm_storeu_si128((__m128i*)&st[0], s0);
    _mm_storeu_si128((__m128i*)&st[4], s1);
}

bool detect_ni()
{
    unsigned a, b, c, d;
    if (!__get_cpuid(1, &a, &b, &c, &d)) return false;
    const bool ssse3 = c & (1u << 9), sse41 = c & (1u << 19);
    if (!__get_cpuid_count(7, 0, &a, &b, &c, &d)) return false;
    const bool sha = b & (1u << 29);
    return ssse3 && sse41 && sha;
}

const bool g_ni = detect_ni();

}  // namespace

bool sha256_has_ni() { return g_ni; }

void sha256_blocks(uint32_t st[8], const uint8_t* data, size_t n)
{
    if (g_ni) blocks_ni(st, data, n);
    else blocks_portable(st, data, n);
}

void HostSha256::init()
{
    static const uint32_t iv[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                                   0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
    memcpy(h, iv, sizeof(h));
    nbuf = 0;
    total = 0;
}

void HostSha256::update(const uint8_t* p, size_t n)
{
    total += n;
    if (nbuf) {
        const size_t take = n < 64 - nbuf ? n : 64 - nbuf;
        memcpy(buf + nbuf, p, take);
        nbuf += (uint32_t)take;
        p += take;
        n -= take;
        if (nbuf < 64) return;
        sha256_blocks(h, buf, 1);
        nbuf = 0;
    }
    const size_t full = n / 64;
    if (full) sha256_blocks(h, p, full);
    p += full * 64;
    n -= full * 64;
    if (n) {
        memcpy(buf, p, n);
        nbuf = (uint32_t)n;
    }
}

void HostSha256::peek(uint8_t out[32]) const
{
    uint32_t st[8];
    memcpy(st, h, sizeof(st));
    uint8_t tail[128];
    memcpy(tail, buf, nbuf);
    size_t len = nbuf;
    tail[len++] = 0x80;
    const size_t padded = (len + 8 <= 64) ? 64 : 128;
    memset(tail + len, 0, padded - len);
    const uint64_t bits = total * 8;
    for (int i = 0; i < 8; ++i) tail[padded - 1 - i] = (uint8_t)(bits >> (8 * i));
    sha256_blocks(st, tail, padded / 64);
    for (int i = 0; i < 8; ++i) {
        out[4 * i] = (uint8_t)(st[i] >> 24);
        out[4 * i + 1] = (uint8_t)(st[i] >> 16);
        out[4 * i + 2] = (uint8_t)(st[i] >> 8);
        out[4 * i + 3] = (uint8_t)st[i];
    }
}

}  // namespace tp

extern "C" {

// SHA-256 of a buffer with the transcript's host hash (absorbed in pieces of
// `chunk` bytes when chunk > 0, to exercise the streaming path).
void tp_sha256(const uint8_t* data, size_t n, size_t chunk, uint8_t out[32])
{
    tp::HostSha256 h;
    h.init();
    if (chunk == 0) chunk = n ? n : 1;
    for (size_t at = 0; at < n; at += chunk) h.update(data + at, n - at < chunk ? n - at : chunk);
    h.peek(out);
}

int tp_sha256_has_ni(void) { return tp::sha256_has_ni() ? 1 : 0; }

}  // extern "C"
