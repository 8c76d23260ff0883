// transcript.hpp - the TruePix Fiat-Shamir transcript byte format, for host
// and device code (TP/libTruePix/fiatshamir.py:27-166, protocol "truepix",
// binary absorb mode).
//
// Absorbed byte stream:
//   prefix      "TruePix," protocol "," "binhash1,"                 (:66-70)
//   int list    0x01 <u32 count> <u64 x count>   (a non-empty list of ints)
//   int         0x02 <u64>
//   Fp2         0x03 <u64 re> <u64 im>
//   str         0x04 <u32 len> <utf-8 bytes>
//   None        0x05
// rand_scalar absorbs the draw counter as a one-element int list, digests the
// running hash and keeps the low 61 bits (the big-endian digest & (2^61 - 1)),
// retrying on the single value p (:144-152).  rand_ext(label) frames each
// coordinate draw with "ext", label, "real"/"imag" (:154-166).
#pragma once
#include "fp61.hpp"
#include "sha256.hpp"

namespace tp {

// H is the running hash: tp::Sha256 (host and device) or tp::HostSha256
// (host only, SHA-NI).
template <class H>
struct FsTranscriptT {
    H sha;
    u64 rnum;

    TP_HD void raw(const void* p, size_t n) { sha.update((const uint8_t*)p, n); }
    TP_HD void le32(uint32_t x)
    {
        uint8_t b[4];
        for (int i = 0; i < 4; ++i) b[i] = (uint8_t)(x >> (8 * i));
        raw(b, 4);
    }
    TP_HD void le64(u64 x)
    {
        uint8_t b[8];
        for (int i = 0; i < 8; ++i) b[i] = (uint8_t)(x >> (8 * i));
        raw(b, 8);
    }
    TP_HD static uint32_t slen(const char* s)
    {
        uint32_t n = 0;
        while (s[n]) ++n;
        return n;
    }

    TP_HD void init(const char* protocol)
    {
        sha.init();
        rnum = 0;
        raw("TruePix,", 8);
        raw(protocol, slen(protocol));
        raw(",", 1);
        raw("binhash1,", 9);
    }

    TP_HD void int_list(const u64* v, uint32_t n)
    {
        const uint8_t t = 0x01;
        raw(&t, 1);
        le32(n);
        for (uint32_t i = 0; i < n; ++i) le64(v[i]);
    }
    TP_HD void item_int(u64 x)
    {
        const uint8_t t = 0x02;
        raw(&t, 1);
        le64(x);
    }
    TP_HD void item_fp2(fp2 x)
    {
        const uint8_t t = 0x03;
        raw(&t, 1);
        le64(x.re);
        le64(x.im);
    }
    TP_HD void item_str(const char* s, uint32_t n)
    {
        const uint8_t t = 0x04;
        raw(&t, 1);
        le32(n);
        raw(s, n);
    }
    TP_HD void item_str(const char* s) { item_str(s, slen(s)); }
    TP_HD void item_none()
    {
        const uint8_t t = 0x05;
        raw(&t, 1);
    }

    TP_HD u64 rand_scalar()
    {
        for (;;) {
            int_list(&rnum, 1);
            ++rnum;
            uint8_t d[32];
            sha.peek(d);
            u64 v = 0;
            for (int i = 24; i < 32; ++i) v = (v << 8) | d[i];
            v &= P;
            if (v < P) return v;
        }
    }

    TP_HD fp2 rand_ext(const char* label, uint32_t n)
    {
        item_str("ext", 3);
        item_str(label, n);
        item_str("real", 4);
        const u64 re = rand_scalar();
        item_str("ext", 3);
        item_str(label, n);
        item_str("imag", 4);
        const u64 im = rand_scalar();
        return mk2(re, im);
    }
};

using FsTranscript = FsTranscriptT<Sha256>;

// Writes "<prefix><k>" into out (no terminator) and returns its length.
TP_HD uint32_t label_num(char* out, const char* prefix, int k)
{
    uint32_t n = 0;
    while (prefix[n]) { out[n] = prefix[n]; ++n; }
    char tmp[12];
    int m = 0;
    unsigned v = (unsigned)k;
    do { tmp[m++] = (char)('0' + v % 10); v /= 10; } while (v);
    while (m) out[n++] = tmp[--m];
    return n;
}

}  // namespace tp
