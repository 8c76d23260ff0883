#!/usr/bin/env python3
"""Writes the golden proof fixtures under tests/golden/ by running the
reference's own pure-Python NIZK prover and verifier
(TP/libTruePix/circuitnizkvec_vole.py) in the build container.

Pinned this way (nothing else fixes them): the transcript byte stream, the
VOLE commitment sequence and the pickled proof bytes (A13-A15), end to end
through the GKR engine.  Run only where /root/reference exists:

    python tests/golden/make_golden_reference.py              # every case
    python tests/golden/make_golden_reference.py --only ml3   # some cases, merged

The reference package is imported from /root/reference with its native bridge
module blocked, so the reference runs its pure-Python arithmetic and the
prebuilt native library it ships is never loaded.  Inputs, VOLE correlations,
manifest and pid are fixed; the fixtures hold the inputs' seeds, the VOLE and
manifest files, and the proof bytes (or their SHA-256 for the large cases).

Two families of cases:
  - the four edit circuits at frame sizes from 5 to 8040 copies (8040 is a
    1080p frame: nCopyBits 13);
  - seeded random layered circuits of 3 and 4 layers over every gate op
    (shaped like TP/libTruePixtests/layer.py:21-94 and circuit.py:21-77, whose
    rand_ckt draws gates the same way), written to multilayer.json with their
    inputs, so the layer loop, the muls/z1_2 hand-off between layers and the
    line projection of CircuitProverNIZK.run (circuitnizkvec_vole.py:387-481)
    are pinned.  Their output layers hold 257..384 gates because the
    reference verifier reads output.txt as 384 values per copy padded to 512
    (circuitnizkvec_vole.py:714-753).  One MUX case records what the
    reference's own verifier says of its own proof (its three MUX definitions
    disagree: arithcircuit.py:38-43, gateprover.py:309-317,
    native_fp2_bridge.py:450-455).
"""
from __future__ import annotations

import argparse
import bz2
import hashlib
import importlib.abc
import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

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
    ("gray3600", "gray", 3600, 106, False),
    ("inv3600", "invert", 3600, 107, False),
    ("crop8040", "crop", 8040, 108, False),
    ("mask8040", "mask", 8040, 109, False),
]

OPS_NO_MUX = ["ADD", "MUL", "SUB", "OR", "XOR", "NOT", "NAND", "NOR", "NXOR", "NAAB", "PASS"]

# (name, seed, input layer (None = private, int = constant), gate counts from the
#  first layer to the output layer, n_copies, ops, muxbits)
ML_CASES = [
    ("ops1", 200, [None] * 20, [300], 6, OPS_NO_MUX, None),
    ("ml2lin", 204, [None] * 20, [40, 300], 6, ["ADD", "SUB", "PASS"], None),
    ("ml3", 201, [None] * 17 + [3, 1, 0], [40, 100, 300], 6, OPS_NO_MUX, None),
    ("ml4", 202, [None] * 9 + [7] + [None] * 2, [16, 8, 64, 384], 3, OPS_NO_MUX, None),
    ("mlmux", 203, [None] * 12, [16, 260], 4, OPS_NO_MUX + ["MUX"] * 4, [1, 0]),
] + [("op_%s" % op.lower(), 210 + k, [None] * 20, [300], 4, [op], None)
     for k, op in enumerate(OPS_NO_MUX)]


def random_layered(seed, input_layer, widths, ops, n_copies, muxlen):
    """A seeded layered circuit [input_layer, layer_1, ..., layer_L] with every
    op of `ops` in every layer (the rest drawn uniformly) and each gate's
    inputs drawn over the previous layer's wires, plus private input values
    in 0..255 for n_copies copies."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parsed = [list(input_layer)]
    n_prev = len(input_layer)
    for w in widths:
        kinds = list(ops) * (w // len(ops)) + [ops[i] for i in rng.integers(0, len(ops), w % len(ops))]
        kinds = [kinds[i] for i in rng.permutation(len(kinds))]
        layer = []
        for op in kinds:
            a, b = (int(x) for x in rng.integers(0, n_prev, 2))
            mx = int(rng.integers(0, muxlen)) if op == "MUX" else None
            layer.append((op, a, b, mx))
        parsed.append(layer)
        n_prev = w
    n_priv = sum(1 for x in input_layer if x is None)
    priv = rng.integers(0, 256, size=(n_copies, n_priv), dtype=np.int64)
    return parsed, priv


def ml_fixture(cases):
    """multilayer.json: the circuits, private inputs and mux bits of ML_CASES."""
    out = []
    for name, seed, inl, widths, n_copies, ops, muxbits in cases:
        parsed, priv = random_layered(seed, inl, widths, ops, n_copies, len(muxbits or [0]))
        out.append({"name": name, "seed": seed, "n_copies": n_copies, "muxbits": muxbits,
                    "input_layer": parsed[0], "layers": [[list(g) for g in l] for l in parsed[1:]],
                    "private": priv.tolist()})
    return out


def parsed_of(entry):
    return [entry["input_layer"]] + [[tuple(g) for g in l] for l in entry["layers"]]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="", help="comma-separated case names (merged into the index)")
    args = ap.parse_args()
    only = set(x for x in args.only.split(",") if x)
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

    ml = ml_fixture(ML_CASES)
    with open(os.path.join(HERE, "multilayer.json"), "w") as fh:
        json.dump({"cases": ml}, fh)

    idx_path = os.path.join(HERE, "golden_proofs.json")
    index = {"pid": PID, "vole_seed": 777, "cases": []}
    if only and os.path.isfile(idx_path):
        with open(idx_path) as fh:
            index = json.load(fh)
        index.setdefault("cases", [])

    todo = []
    for name, circ, n_copies, seed, full in CASES:
        parsed = circuits.circuit(circ)
        priv = random_private_inputs(circ, parsed, n_copies, seed=seed)
        todo.append((name, dict(circuit=circ, input_seed=seed), parsed, priv, n_copies, None, full, seed))
    for e in ml:
        todo.append((e["name"], dict(circuit="multilayer", multilayer=e["name"]), parsed_of(e),
                     np.array(e["private"], dtype=np.int64), e["n_copies"], e["muxbits"], True, e["seed"]))

    for name, meta, parsed, priv, n_copies, muxbits, full, seed in todo:
        if only and name not in only:
            continue
        rows = full_input_rows(parsed, priv).tolist()
        manifest = make_manifest(n_copies, seed=seed)
        man_path = os.path.join(HERE, "manifest_%s.json" % name)
        write_manifest(man_path, manifest)
        work = tempfile.mkdtemp(prefix="tpgold_")
        cwd = os.getcwd()
        accepted = None
        try:
            shutil.copy(vole_p, work)
            shutil.copy(vole_v, work)
            shutil.copy(man_path, os.path.join(work, "orion_commit.json"))
            os.chdir(work)
            (_, prv) = ref_nizk.CircuitProverNIZK.from_pws(parsed, n_copies)
            t0 = time.time()
            try:
                proof = prv.run([list(r) for r in rows], muxbits)
            except TypeError as exc:
                if muxbits is None:
                    raise
                # the reference's NIZK prover cannot prove MUX gates: recorded
                raises = "%s: %s" % (type(exc).__name__, exc)
                print("%-10s reference prover raised %s" % (name, raises), flush=True)
                entry = dict(name=name, n_copies=n_copies, manifest=os.path.basename(man_path),
                             muxbits=muxbits, reference_prover_raises=raises)
                entry.update(meta)
                index["cases"] = [c for c in index["cases"] if c["name"] != name] + [entry]
                with open(idx_path, "w") as fh:
                    json.dump(index, fh, indent=1)
                continue
            t_prove = time.time() - t0
            outs = [list(o[:384]) for o in prv.prover.ckt_outputs]
            with open("output.txt", "w") as fh:
                for o in outs:
                    o = [int(v) for v in o] + [0] * (384 - len(o))
                    fh.write(" ".join(str(v) for v in o) + "\n")
            (_, ver) = ref_nizk.CircuitVerifierNIZK.from_pws(parsed, n_copies)
            try:
                pending = ver.run(proof)
                invals = []
                for r in rows:
                    invals.extend(list(r) + [0] * (2 ** ver.nInBits - len(r)))
                invals += [0] * (2 ** (ver.nInBits + ver.nCopyBits) - len(invals))
                value = VerifierIOMLExt(pending.input_point).compute(invals)
                accepted = bool(ver.finish_input(value))
            except Exception as exc:   # recorded, not hidden: only the MUX case may land here
                print("  reference verifier raised %s: %s" % (type(exc).__name__, exc), flush=True)
                accepted, pending, value = False, None, None
        finally:
            os.chdir(cwd)
            shutil.rmtree(work, ignore_errors=True)
        digest = hashlib.sha256(proof).hexdigest()
        entry = dict(name=name, n_copies=n_copies, manifest=os.path.basename(man_path),
                     proof_sha256=digest, proof_len=len(proof),
                     reference_prove_seconds=round(t_prove, 3),
                     reference_verifier_accepts=accepted)
        entry.update(meta)
        if muxbits is not None:
            entry["muxbits"] = muxbits
        if pending is not None:
            entry["input_value"] = [value.real, value.imag]
            entry["input_point"] = [[e.real, e.imag] for e in pending.input_point]
        if full:
            fname = "proof_%s.bin.bz2" % name
            with open(os.path.join(HERE, fname), "wb") as fh:
                fh.write(bz2.compress(proof))
            entry["proof_file"] = fname
        index["cases"] = [c for c in index["cases"] if c["name"] != name] + [entry]
        print("%-10s %6d copies  proof %8d B  sha256 %s  (reference prove %.1f s, verifier %s)"
              % (name, n_copies, len(proof), digest[:16], t_prove, accepted), flush=True)
        with open(idx_path, "w") as fh:
            json.dump(index, fh, indent=1)


if __name__ == "__main__":
    main()
