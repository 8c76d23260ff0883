// relay.hpp - round results to the host and challenges back through mapped
// host memory as tagged words (no fence, no separate sequence word), for
// kernels that stay resident across the rounds of a sumcheck whose
// challenges the host's transcript draws.
//
// Every word carries (sequence number mod 8) << 61 above its canonical
// (61-bit) value, so one read of a word is both the check and the value.
// The host empties a slot (all ones: value bits p, never canonical) once it
// has read it; answers overwrite one another.  A kernel's waits are bounded
// by the wall clock; an expired wait stores kSeqAbort to the slot's sequence
// word and the kernel ends.
#pragma once
#include "fp61.hpp"

namespace tp {

constexpr unsigned kSeqAbort = 0xffffffffu;   // a kernel's bounded wait expired
constexpr unsigned kSeqHostAbort = 0xfffffffdu;   // the host gave up: waiting launches leave
constexpr u64 kSlotEmpty = ~0ull;   // value bits p under tag 7: no round's word

// A round's sums to the host's mapped slot: every word carries (seq mod 8)
// << 61 above its canonical value, so the host checks each word itself and
// no fence has to order the values before a sequence word (a system fence
// cost about 0.4 us a round).  The host empties the slot (all ones) once it
// has read it, before it answers or launches again, so a stale word never
// passes for a later round.  Aborts still go to the sequence word.
template <int N>
__device__ __forceinline__ void publish_tagged(fp2* host_vals, const fp2 (&v)[N], unsigned seq)
{
    const u64 tag = (u64)(seq & 7) << 61;
    u64* w = reinterpret_cast<u64*>(host_vals);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        __hip_atomic_store(w + 2 * k, v[k].re | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(w + 2 * k + 1, v[k].im | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// The host's answer to the round published as `want`, from the mapped answer
// slot: the host writes the N words (N / 2 challenges) with (want mod 8) << 61
// above the canonical value, so one read of the slot is both the check and
// the value (a separate sequence word cost a second PCIe round trip).  A word
// of another round, or the empty mark (all ones: value bits p), does not
// match.  The host empties the slot when a prove starts and each answer
// overwrites the one before; answered rounds in a prove are never 8 sequence
// numbers apart with no answer between, so a stale answer never matches.
// Returns 0 if the wait expires.
template <int N>
__device__ __forceinline__ int take_answer_n(const u64* answer, unsigned want, long long t0,
                                             long long spin_ticks, u64 (&rv)[N])
{
    const u64 tag = (u64)(want & 7) << 61;
    for (;;) {
        u64 a[N];
        bool all = true;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            a[i] = __hip_atomic_load(answer + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            all = all && (a[i] & ~P) == tag && (a[i] & P) != P;
        }
        if (all) {
#pragma unroll
            for (int i = 0; i < N; ++i) rv[i] = a[i] & P;
            return 1;
        }
        if (wall_clock64() - t0 > spin_ticks) return 0;
        __builtin_amdgcn_s_sleep(1);
    }
}

__device__ __forceinline__ int take_answer(const u64* answer, unsigned want, long long t0,
                                           long long spin_ticks, u64 (&rv)[2])
{
    return take_answer_n<2>(answer, want, t0, spin_ticks, rv);
}

}  // namespace tp
