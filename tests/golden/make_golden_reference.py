#!/usr/bin/env python3
"""Writes the golden proof fixtures under tests/golden/ by running the
reference's own pure-Python NIZK prover and verifier
(TP/libTruePix/circuitnizkvec_vole.py) in the build container.

Pinned this way (nothing else fixes them): the transcript byte stream, the
VOLE commitment sequence and the pickled proof bytes (A13-A15), end to end
through the GKR engine.  Run only where /root/reference exists:

    python tests/golden/make_golden_reference.py

The reference package is imported from /root/reference with its native bridge
module blocked, so the reference runs its pure-Python arithmetic and the
prebuilt native library it ships is never loaded.  Inputs, VOLE correlations,
manifest and pid are fixed; the fixtures hold the inputs' seeds, the VOLE and
manifest files, and the proof bytes (or their SHA-256 for the large case).
"""
from __future__ import annotations

import bz2
import hashlib
import importlib.abc
import json
import os
import shutil
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/TruePix-main"
PID = 4242

sys.dont_write_bytecode = True
sys.path.insert(0, ROOT)


class _BlockNative(importlib.abc.MetaPathFinder):
    """Makes `import libTruePix.native_fp2_bridge` fail, which the reference
    treats as 'no native library' (pure-Python path)."""

    def find_spec(self, name, path=None, target=None):
        if name == "libTruePix.native_fp2_bridge":
            raise ImportError("native bridge disabled for golden generation")
        return None


# (name, circuit, n_copies, input seed, store full proof bytes)
CASES = [
    ("gray16", "gray", 16, 101, True),
    ("inv8", "invert", 8, 102, True),
    ("crop12", "crop", 12, 103, True),
    ("mask5", "mask", 5, 104, True),
    ("gray1350", "gray", 1350, 105, False),
]


def main():
    if not os.path.isdir(REF):
        raise SystemExit("reference tree not present")
    sys.meta_path.insert(0, _BlockNative())
    sys.path.insert(0, REF)
    from libTruePix.defs import Defs
    Defs.configure()
    import libTruePix.circuitnizkvec_vole as ref_nizk
    from libTruePix.iomlext import VerifierIOMLExt

    from truepix_amd import circuits
    from truepix_amd.fp2_link import full_input_rows, make_manifest, random_private_inputs, write_manifest
    from truepix_amd.vole import ExtVOLEGenerator

    class _OsShim:
        path = os.path

        @staticmethod
        def getpid():
            return PID
    ref_nizk.os = _OsShim()          # fixed pid inside the reference prover

    vole_p = os.path.join(HERE, "vole_ext_prover_%d.json" % PID)
    vole_v = os.path.join(HERE, "vole_ext_verifier_%d.json" % PID)
    gen = ExtVOLEGenerator(seed=777)
    gen.generate_deltas()
    gen.generate_vole_instances(500)
    gen.save_to_files(vole_p, vole_v)

    index = {"pid": PID, "vole_seed": 777, "cases": []}
    for name, circ, n_copies, seed, full in CASES:
        parsed = circuits.circuit(circ)
        priv = random_private_inputs(circ, parsed, n_copies, seed=seed)
        rows = full_input_rows(parsed, priv).tolist()
        manifest = make_manifest(n_copies, seed=seed)
        man_path = os.path.join(HERE, "manifest_%s.json" % name)
        write_manifest(man_path, manifest)
        work = tempfile.mkdtemp(prefix="tpgold_")
        cwd = os.getcwd()
        try:
            shutil.copy(vole_p, work)
            shutil.copy(vole_v, work)
            shutil.copy(man_path, os.path.join(work, "orion_commit.json"))
            os.chdir(work)
            (_, prv) = ref_nizk.CircuitProverNIZK.from_pws(parsed, n_copies)
            t0 = time.time()
            proof = prv.run([list(r) for r in rows])
            t_prove = time.time() - t0
            outs = [list(o[:384]) for o in prv.prover.ckt_outputs]
            with open("output.txt", "w") as fh:
                for o in outs:
                    fh.write(" ".join(str(int(v)) for v in o) + "\n")
            (_, ver) = ref_nizk.CircuitVerifierNIZK.from_pws(parsed, n_copies)
            pending = ver.run(proof)
            invals = []
            for r in rows:
                invals.extend(list(r) + [0] * (2 ** ver.nInBits - len(r)))
            invals += [0] * (2 ** (ver.nInBits + ver.nCopyBits) - len(invals))
            value = VerifierIOMLExt(pending.input_point).compute(invals)
            assert ver.finish_input(value)
        finally:
            os.chdir(cwd)
            shutil.rmtree(work, ignore_errors=True)
        digest = hashlib.sha256(proof).hexdigest()
        entry = {"name": name, "circuit": circ, "n_copies": n_copies, "input_seed": seed,
                 "manifest": os.path.basename(man_path), "proof_sha256": digest,
                 "proof_len": len(proof), "reference_prove_seconds": round(t_prove, 3),
                 "input_value": [value.real, value.imag],
                 "input_point": [[e.real, e.imag] for e in pending.input_point]}
        if full:
            fname = "proof_%s.bin.bz2" % name
            with open(os.path.join(HERE, fname), "wb") as fh:
                fh.write(bz2.compress(proof))
            entry["proof_file"] = fname
        index["cases"].append(entry)
        print("%-10s %6d copies  proof %8d B  sha256 %s  (reference prove %.1f s)"
              % (name, n_copies, len(proof), digest[:16], t_prove), flush=True)
    with open(os.path.join(HERE, "golden_proofs.json"), "w") as fh:
        json.dump(index, fh, indent=1)


if __name__ == "__main__":
    main()
