#!/usr/bin/env python3
"""Writes tests/golden/circuits.json: SHA-256 of each of the four TruePix edit
circuits as parsed from the reference's PWS files (TP/pws/*.pws) with
truepix_amd.circuits.parse_pws_text.  The restated builders in
truepix_amd/circuits.py are checked against these digests by
tests/test_protocol_cpu.py, so the builders are pinned to the PWS text.
Run only where /root/reference exists."""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from truepix_amd import circuits  # noqa: E402

PWS = "/root/reference/TruePix-main/pws"
FILES = {"gray": "gray.pws", "invert": "inv.pws", "crop": "Crop.pws", "mask": "mask.pws"}


def digest(parsed):
    return hashlib.sha256(repr(parsed).encode()).hexdigest()


def main():
    out = {}
    for name, f in FILES.items():
        parsed = circuits.parse_pws_file(os.path.join(PWS, f))
        out[name] = {"pws": "TP/pws/" + f, "sha256": digest(parsed),
                     "n_inputs": len(parsed[0]), "n_gates": len(parsed[1])}
        print(name, out[name])
    with open(os.path.join(HERE, "circuits.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
