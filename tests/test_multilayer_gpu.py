"""GPU parity of the Python-driven device prover (gkr.ProverClass) on layered
circuits other than the four one-layer edit circuits: seeded random circuits
of 1 to 4 layers over every gate op but MUX, proved by the reference's own
pure-Python prover (TP/libTruePix/circuitnizkvec_vole.py:265-501) into
tests/golden/ by tests/golden/make_golden_reference.py.

Pinned: the layer loop and the hand-off between layers (muls, z1_2, the line
projection of the second-to-last layer, circuitnizkvec_vole.py:387-481), the
early/late rounds of every gate type and the wiring predicate of each.

The reference's verifier rejects the reference's own proofs for circuits with
OR, XOR, NOT, NAND, NOR or NXOR gates: its committed gate functions take the
product of two MAC values where the prover commits the product of the values
(TP/libTruePix/gateprover_vole.py:363-374 against :370-374), and constants
enter unscaled by the VOLE key.  Proof bytes are compared for every case, and
this package's verifier must reach the reference verifier's verdict.
For MUX gates the reference's NIZK prover raises (its eval_mlext indexes
per-layer mux bits that the prover never sets, circuitverifier.py:276), so the
device prover refuses MUX circuits with an error, as the reference does."""
import bz2
import json
import os
import shutil

import numpy as np
import pytest

from truepix_amd import gkr
from truepix_amd.fp2_link import full_input_rows

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _index():
    with open(os.path.join(GOLD, "golden_proofs.json")) as fh:
        idx = json.load(fh)
    with open(os.path.join(GOLD, "multilayer.json")) as fh:
        circs = {c["name"]: c for c in json.load(fh)["cases"]}
    return idx["pid"], [e for e in idx["cases"] if e.get("circuit") == "multilayer"], circs


PID, ENTRIES, CIRCS = _index()
NAMES = [e["name"] for e in ENTRIES if e.get("muxbits") is None]


def _parsed(c):
    return [c["input_layer"]] + [[tuple(g) for g in lay] for lay in c["layers"]]


def _setup(tmp_path, entry):
    for f in ("vole_ext_prover_%d.json" % PID, "vole_ext_verifier_%d.json" % PID):
        shutil.copy(os.path.join(GOLD, f), tmp_path / f)
    shutil.copy(os.path.join(GOLD, entry["manifest"]), tmp_path / "orion_commit.json")


@pytest.mark.parametrize("name", NAMES)
def test_layered_proof_bytes_match_reference(bridge, torch_cuda, tmp_path, name):
    entry = next(e for e in ENTRIES if e["name"] == name)
    circ = CIRCS[name]
    parsed = _parsed(circ)
    _setup(tmp_path, entry)
    rows = full_input_rows(parsed, np.array(circ["private"], dtype=np.int64))
    _, prv = gkr.ProverClass.from_pws(parsed, entry["n_copies"])
    prv.pid = PID
    prv.vole_dir = str(tmp_path)
    prv.manifest_path = str(tmp_path / "orion_commit.json")
    print("layered case", name, [len(l) for l in circ["layers"]], flush=True)
    proof = prv.run(rows.tolist())
    with open(os.path.join(GOLD, entry["proof_file"]), "rb") as fh:
        gold = bz2.decompress(fh.read())
    assert len(proof) == len(gold)
    assert proof == gold
    # the verifier here reaches the reference verifier's verdict on the same bytes
    outs = prv.prover.ckt_outputs[:entry["n_copies"], :384]
    with open(tmp_path / "output.txt", "w") as fh:
        for row in outs.tolist():
            row = [int(v) for v in row] + [0] * (384 - len(row))
            fh.write(" ".join(str(v) for v in row) + "\n")
    _, ver = gkr.VerifierClass.from_pws(parsed, entry["n_copies"])
    ver.vole_dir = str(tmp_path)
    ver.manifest_path = str(tmp_path / "orion_commit.json")
    ver.output_path = str(tmp_path / "output.txt")
    if entry["reference_verifier_accepts"]:
        pending = ver.run(proof)
        assert [[e.real, e.imag] for e in pending.input_point] == entry["input_point"]
        assert (prv.final_input_value.real, prv.final_input_value.imag) == tuple(entry["input_value"])
        assert ver.finish_input(prv.final_input_value)
    else:
        with pytest.raises(ValueError, match="Verification"):
            ver.run(proof)


def test_mux_circuit_is_refused_as_the_reference_does(bridge, tmp_path):
    entry = next(e for e in ENTRIES if e.get("muxbits") is not None)
    circ = CIRCS[entry["name"]]
    parsed = _parsed(circ)
    _setup(tmp_path, entry)
    rows = full_input_rows(parsed, np.array(circ["private"], dtype=np.int64))
    _, prv = gkr.ProverClass.from_pws(parsed, entry["n_copies"])
    prv.pid = PID
    prv.vole_dir = str(tmp_path)
    prv.manifest_path = str(tmp_path / "orion_commit.json")
    with pytest.raises(ValueError, match="MUX"):
        prv.run(rows.tolist(), entry["muxbits"])
