#!/usr/bin/env python3
"""Writes tests/golden/fft_gkr_ref.json from the reference's own fft_gkr
(OZ/src/VPD/fft_circuit_GKR.cpp compiled in place by `make -C oracle ref`,
driven through its layer functions by oracle/ref_harness_fftgkr.cpp):
for lg_size 1..12, the pass flag, the proof size, the claim after every layer
with the final alpha / beta, and SHA-256 digests of the circuit's last IFFT
layer, its 1/n-scaled layer and the 64 sums.  Run where /root/reference
exists:

    make -C oracle ref && python tests/golden/make_golden_fftgkr.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def digests(layers, lg):
    n = 1 << lg
    raw = layers.tobytes()
    return {"ifft_out": hashlib.sha256(raw[:16 * n]).hexdigest(),
            "scaled": hashlib.sha256(raw[16 * n:32 * n]).hexdigest(),
            "sums": hashlib.sha256(raw[32 * n:32 * n + 64 * 16]).hexdigest()}


def main():
    from oracle.oracle import RefFftGkr
    if not RefFftGkr.available():
        raise SystemExit("build the reference harness first: make -C oracle ref")
    ref = RefFftGkr()
    out = {"source": "OZ/src/VPD/fft_circuit_GKR.cpp (engage_gkr order), glibc rand() seed 1",
           "cases": []}
    for lg in range(1, 13):
        r = ref.run(lg)
        out["cases"].append({"lg_size": lg, "ok": int(r["ok"]), "proof_size": r["proof_size"],
                             "claims": [int(x) for x in r["claims"]],
                             "layer_sha256": digests(r["layers"], lg)})
        print(lg, r["ok"], r["proof_size"], flush=True)
    with open(os.path.join(HERE, "fft_gkr_ref.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
