// CPU check of the early-claim term table (truepix_amd/csrc/claim_terms.hpp):
// for random layers over all 13 gate types, random z1 and random wire values,
// the claim of one pair of copies at t = 0, 1, -1, 2 summed gate by gate
// (sum_g z1[g] gate_g(V_t[in0], V_t[in1]), V_-1 = 2 V_0 - V_1, V_2 = 2 V_1 - V_0,
// TP/native_fp2/native_fp2_core.hpp:200-248) equals the same claim over the
// terms: non-affine gates as themselves, each wire's summed weight times its
// value at t = 0 and 1, the constant, and the affine part's values at -1 and 2
// by linearity.  Prints each failure; exit status 1 on any.
#include "claim_terms.hpp"
#include "fp61.hpp"
#include <cstdio>
#include <vector>

using namespace tp;

static int fails = 0;
static u64 rng_state = 0x243f6a8885a308d3ull;
static u64 rnd() {
    rng_state ^= rng_state << 13; rng_state ^= rng_state >> 7; rng_state ^= rng_state << 17;
    return rng_state;
}
static fp2 rfp2() { return mk2(rnd() % P, rnd() % P); }

int main()
{
    int layers = 0;
    for (int trial = 0; trial < 300; ++trial) {
        const int n_wires = 1 + (int)(rnd() % 40), n_gates = 1 + (int)(rnd() % 60);
        std::vector<int> type(n_gates), in0(n_gates), in1(n_gates);
        for (int g = 0; g < n_gates; ++g) {
            // bias towards few wires so that gates share them
            type[g] = (int)(rnd() % 13);
            in0[g] = (int)(rnd() % n_wires);
            in1[g] = trial % 3 == 0 ? in0[g] : (int)(rnd() % n_wires);
        }
        std::vector<fp2> z(n_gates), v0(n_wires), v1(n_wires);
        for (auto& x : z) x = rfp2();
        for (int w = 0; w < n_wires; ++w) { v0[w] = rfp2(); v1[w] = rfp2(); }
        auto at = [&](const std::vector<fp2>& a, const std::vector<fp2>& b, int w, int t) {
            switch (t) {
            case 0: return a[w];
            case 1: return b[w];
            case 2: return sub2(dbl2(a[w]), b[w]);
            default: return sub2(dbl2(b[w]), a[w]);
            }
        };
        fp2 want[4], got[4];
        for (int t = 0; t < 4; ++t) {
            want[t] = zero2();
            for (int g = 0; g < n_gates; ++g)
                want[t] = add2(want[t], mul2(gatefn(type[g], at(v0, v1, in0[g], t),
                                                    at(v0, v1, in1[g], t)), z[g]));
        }
        const ClaimTerms T = build_claim_terms(n_gates, type.data(), in0.data(), in1.data());
        fp2 l0 = zero2(), l1 = zero2();
        for (int t = 0; t < 4; ++t) got[t] = zero2();
        int n_src = 0;
        for (int k = 0; k < T.n_terms(); ++k) {
            fp2 c = zero2();
            for (int e = T.off[k]; e < T.off[k + 1]; ++e) {
                c = T.neg[e] ? sub2(c, z[T.src[e]]) : add2(c, z[T.src[e]]);
                ++n_src;
            }
            if (T.type[k] == kTermWire) {
                l0 = add2(l0, mul2(v0[T.in0[k]], c));
                l1 = add2(l1, mul2(v1[T.in0[k]], c));
            } else if (T.type[k] == kTermConst) {
                l0 = add2(l0, c);
                l1 = add2(l1, c);
            } else {
                for (int t = 0; t < 4; ++t)
                    got[t] = add2(got[t], mul2(gatefn(T.type[k], at(v0, v1, T.in0[k], t),
                                                      at(v0, v1, T.in1[k], t)), c));
            }
        }
        got[0] = add2(got[0], l0);
        got[1] = add2(got[1], l1);
        got[2] = add2(got[2], sub2(dbl2(l0), l1));
        got[3] = add2(got[3], sub2(dbl2(l1), l0));
        for (int t = 0; t < 4; ++t)
            if (!eq2(got[t], want[t]) && ++fails <= 20)
                std::printf("trial %d: claim at point %d differs\n", trial, t);
        if ((int)T.off.size() != T.n_terms() + 1 || T.off.back() != (int)T.src.size() ||
            n_src < n_gates)
            if (++fails <= 20) std::printf("trial %d: malformed term table\n", trial);
        ++layers;
    }
    std::printf("%d layers, %d failures\n", layers, fails);
    return fails ? 1 : 0;
}
