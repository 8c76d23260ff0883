// virgo.hpp - the Virgo code-switching proof that closes a NIZK Orion
// opening (OZ/src/VPD/linearPC.cpp:875-905): the queried entries of the
// combined codeword are the expander code of the combined message, proved
// layer by layer over the encoding circuit of generate_circuit
// (linearPC.cpp:477-669) by the Virgo GKR prover (OZ/src/linear_gkr/prover.cpp)
// and checked by its verifier (OZ/src/linear_gkr/verifier.cpp:1031-1211).
#pragma once
#include <functional>
#include <vector>

#include <hip/hip_runtime.h>

#include "fp61.hpp"

namespace tp {
namespace virgo {

// One expander graph of the code, resident on the device, forward and
// degree-major: edge slot e of left vertex i goes to right vertex
// to[e * L + i] with weight w[e * L + i].
struct Graph {
    long long L = 0, R = 0;
    int deg = 0;
    const int* to = nullptr;
    const fp2* w = nullptr;
};

// The code of one message length n: C[d] and D[d] for every depth d whose
// input is longer than the distance threshold (13); cw is the codeword length.
struct Code {
    long long n = 0, cw = 0;
    std::vector<Graph> C, D;
};

struct Stats {
    double ms = 0;      // device time of the proof's launches
    int launches = 0;
    int layers = 0;     // sumcheck layers
};

// Proves and checks the code-switching claim for the combined codeword d_cc
// (cw entries on the device; its first n entries are the combined message)
// at the positions q (query_count draws, sorted here as generate_circuit
// sorts them).  squeeze() yields the transcript's next u64, positioned after
// the "combined-mt" absorb.  words receives the transcript (layout in
// truepix_hip.h, tp_orion_open_nizk_virgo), accepted the verifier's verdict.
bool prove(const Code& code, const fp2* d_cc, std::vector<long long> q,
           const std::function<u64()>& squeeze, hipStream_t st, std::vector<u64>& words,
           bool& accepted, Stats* stats);

}  // namespace virgo
}  // namespace tp
