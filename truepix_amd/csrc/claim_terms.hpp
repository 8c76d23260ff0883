// claim_terms.hpp - the term table of the early-round claim (host side).
//
// The claim sums z1[g] * gate_g(V[in0_g], V[in1_g]) over the gates of a layer
// at four points per pair of copies (TP/native_fp2/native_fp2_core.hpp:200-248).
// A gate that is affine in its inputs, a0 x + a1 y + c with a0, a1, c in
// {-1, 0, 1} (ADD, SUB, MUX-left, MUX-right, NOT, PASS), adds z1[g] a0 to a
// weight on wire in0, z1[g] a1 to one on wire in1 and z1[g] c to a constant,
// for the whole early phase (z1 is fixed while the copies fold).  So the
// claim runs over terms: each non-affine gate as itself, each wire read by an
// affine gate once with its summed weight, and one constant term.  The gray
// circuit's 384 PASS gates become 257 wire terms, mask's become 3.
//
// The table is the circuit's; the coefficients depend on z1 and are formed on
// the device (body::term_coeffs) from the contributions listed here (CSR:
// term k sums z1[src[e]], negated where neg[e], for e in [off[k], off[k+1])).
#pragma once
#include <map>
#include <vector>

namespace tp {

constexpr int kTermWire = 13;    // the affine gates' weight on wire in0
constexpr int kTermConst = 14;   // the affine gates' constant

struct ClaimTerms {
    std::vector<int> type, in0, in1;   // per term (in1 unused for wire / constant terms)
    std::vector<int> off, src, neg;    // contributions, CSR over terms
    int n_terms() const { return (int)type.size(); }
};

// The affine form of a gate type (fp61.hpp gatefn): false if it is not affine.
inline bool affine_gate(int t, int& a0, int& a1, int& c)
{
    a0 = a1 = c = 0;
    switch (t) {
    case 1: a0 = 1; a1 = 1; return true;    // ADD
    case 2: a0 = 1; a1 = -1; return true;   // SUB
    case 3: a0 = 1; return true;            // MUX left
    case 4: a1 = 1; return true;            // MUX right
    case 7: a0 = -1; c = 1; return true;    // NOT
    case 12: a0 = 1; return true;           // PASS
    case 0: case 5: case 6: case 8: case 9: case 10: case 11: return false;
    default: a0 = 1; return true;           // unknown types pass x, as gatefn does
    }
}

inline ClaimTerms build_claim_terms(int n_gates, const int* type, const int* in0, const int* in1)
{
    ClaimTerms T;
    std::map<int, std::vector<std::pair<int, int>>> wires;   // wire -> (gate, negate)
    std::vector<std::pair<int, int>> cst;
    T.off.push_back(0);
    for (int g = 0; g < n_gates; ++g) {
        int a0, a1, c;
        if (!affine_gate(type[g], a0, a1, c)) {
            T.type.push_back(type[g]);
            T.in0.push_back(in0[g]);
            T.in1.push_back(in1[g]);
            T.src.push_back(g);
            T.neg.push_back(0);
            T.off.push_back((int)T.src.size());
            continue;
        }
        if (a0) wires[in0[g]].push_back({g, a0 < 0});
        if (a1) wires[in1[g]].push_back({g, a1 < 0});
        if (c) cst.push_back({g, c < 0});
    }
    for (auto& w : wires) {
        T.type.push_back(kTermWire);
        T.in0.push_back(w.first);
        T.in1.push_back(w.first);
        for (auto& e : w.second) {
            T.src.push_back(e.first);
            T.neg.push_back(e.second);
        }
        T.off.push_back((int)T.src.size());
    }
    if (!cst.empty()) {
        T.type.push_back(kTermConst);
        T.in0.push_back(0);
        T.in1.push_back(0);
        for (auto& e : cst) {
            T.src.push_back(e.first);
            T.neg.push_back(e.second);
        }
        T.off.push_back((int)T.src.size());
    }
    return T;
}

}  // namespace tp
