// glibc_rand.hpp - glibc's rand() (the TYPE_3 additive generator, the state
// srand(seed) leaves), for host code that must draw the same numbers as the
// reference programs: fft_gkr's challenges (OZ/src/VPD/fft_circuit_GKR.cpp,
// seed 1) and the Orion expander graphs (OZ/include/linear_code/expanders.h,
// glibc's initial state).  Restated from the generator's published
// definition: r[0] = seed, r[i] = 16807 r[i-1] mod (2^31 - 1) (i < 31),
// r[i] = r[i-31] (i < 34), then r[k] = r[k-31] + r[k-3] mod 2^32; the output
// is r[k] >> 1 from k = 344 on.
#pragma once
#include "fp61.hpp"

#include <stdint.h>

namespace tp {

struct GlibcRand {
    uint32_t r[34];
    long k = 34;
    explicit GlibcRand(uint32_t seed)
    {
        int32_t s[34];
        s[0] = (int32_t)(seed ? seed : 1);
        for (int i = 1; i < 31; ++i) {
            long long t = (16807LL * s[i - 1]) % 2147483647LL;
            if (t < 0) t += 2147483647LL;
            s[i] = (int32_t)t;
        }
        for (int i = 31; i < 34; ++i) s[i] = s[i - 31];
        for (int i = 0; i < 34; ++i) r[i] = (uint32_t)s[i];
        for (int i = 0; i < 310; ++i) step();
    }
    uint32_t step()
    {
        const uint32_t v = r[(k - 31) % 34] + r[(k - 3) % 34];
        r[k % 34] = v;
        ++k;
        return v;
    }
    uint32_t next() { return step() >> 1; }
    // prime_field::random() (OZ/src/linear_gkr/prime_field.cpp:35-41): real, then imag
    fp2 fe()
    {
        const u64 re = next() % P;
        const u64 im = next() % P;
        return mk2(re, im);
    }
};

}  // namespace tp
