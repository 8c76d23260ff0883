"""End-to-end GPU parity of the truepix GKR prover: for fixed inputs, VOLE
correlations, manifest and pid, the proof bytes equal the reference's
(golden fixtures from the reference's own prover, which its verifier
accepted), and this package's verifier accepts them and rejects tampering."""
import bz2
import hashlib
import json
import os
import pickle
import shutil

import numpy as np
import pytest

from truepix_amd import circuits, gkr
from truepix_amd.extfield import Fp2, proof_pickle_names
from truepix_amd.fiatshamir import FiatShamir
from truepix_amd.fp2_link import full_input_rows, random_private_inputs

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cases():
    with open(os.path.join(GOLD, "golden_proofs.json")) as fh:
        idx = json.load(fh)
    return idx["pid"], idx["cases"]


def setup_dir(tmp_path, pid, entry):
    for f in ("vole_ext_prover_%d.json" % pid, "vole_ext_verifier_%d.json" % pid):
        shutil.copy(os.path.join(GOLD, f), tmp_path / f)
    shutil.copy(os.path.join(GOLD, entry["manifest"]), tmp_path / "orion_commit.json")


def prove(tmp_path, pid, entry):
    parsed = circuits.circuit(entry["circuit"])
    priv = random_private_inputs(entry["circuit"], parsed, entry["n_copies"], seed=entry["input_seed"])
    rows = full_input_rows(parsed, priv)
    _, prv = gkr.ProverClass.from_pws(parsed, entry["n_copies"])
    prv.pid = pid
    prv.vole_dir = str(tmp_path)
    prv.manifest_path = str(tmp_path / "orion_commit.json")
    proof = prv.run(rows.tolist())
    return parsed, prv, proof


def write_outputs(tmp_path, prv, n_copies):
    outs = prv.prover.ckt_outputs[:n_copies, :384]
    with open(tmp_path / "output.txt", "w") as fh:
        for row in outs.tolist():
            fh.write(" ".join(str(v) for v in row) + "\n")


@pytest.mark.parametrize("name", ["gray16", "inv8", "crop12", "mask5", "gray1350", "inv3600",
                                  "mask8040"])
def test_proof_bytes_match_reference(bridge, torch_cuda, tmp_path, name):
    pid, entries = cases()
    entry = next(e for e in entries if e["name"] == name)
    setup_dir(tmp_path, pid, entry)
    parsed, prv, proof = prove(tmp_path, pid, entry)
    assert len(proof) == entry["proof_len"]
    if "proof_file" in entry:
        with open(os.path.join(GOLD, entry["proof_file"]), "rb") as fh:
            gold = bz2.decompress(fh.read())
        assert proof == gold
    assert hashlib.sha256(proof).hexdigest() == entry["proof_sha256"]
    # the device-computed input MLE equals the reference's value at the point
    assert (prv.final_input_value.real, prv.final_input_value.imag) == tuple(entry["input_value"])
    # this package's verifier accepts it
    write_outputs(tmp_path, prv, entry["n_copies"])
    _, ver = gkr.VerifierClass.from_pws(parsed, entry["n_copies"])
    ver.vole_dir = str(tmp_path)
    ver.manifest_path = str(tmp_path / "orion_commit.json")
    ver.output_path = str(tmp_path / "output.txt")
    pending = ver.run(proof)
    assert [[e.real, e.imag] for e in pending.input_point] == entry["input_point"]
    assert ver.finish_input(Fp2(*entry["input_value"]))


def _tamper(proof, which):
    (q, io, trans, ndb, rvs, rve, wdiv) = FiatShamir.unpack_proof(proof)
    trans = list(trans)
    item = list(trans[which])
    v = item[0]
    if isinstance(v, tuple):
        item[0] = (v[0] + 1,) + tuple(v[1:])
    else:
        item[0] = v + 1
    trans[which] = item
    with proof_pickle_names():
        return pickle.dumps((q, list(io), trans, ndb, rvs, rve, wdiv, 4), -1)


@pytest.mark.parametrize("which", [2, 5, 20, -3, -1])
def test_verifier_rejects_tampered_proof(bridge, torch_cuda, tmp_path, which):
    pid, entries = cases()
    entry = next(e for e in entries if e["name"] == "gray16")
    setup_dir(tmp_path, pid, entry)
    parsed, prv, proof = prove(tmp_path, pid, entry)
    write_outputs(tmp_path, prv, entry["n_copies"])
    bad = _tamper(proof, which)
    _, ver = gkr.VerifierClass.from_pws(parsed, entry["n_copies"])
    ver.vole_dir = str(tmp_path)
    ver.manifest_path = str(tmp_path / "orion_commit.json")
    ver.output_path = str(tmp_path / "output.txt")
    with pytest.raises(ValueError):
        ver.run(bad)
        ver.finish_input(Fp2(*entry["input_value"]))


def test_verifier_rejects_wrong_outputs(bridge, torch_cuda, tmp_path):
    pid, entries = cases()
    entry = next(e for e in entries if e["name"] == "inv8")
    setup_dir(tmp_path, pid, entry)
    parsed, prv, proof = prove(tmp_path, pid, entry)
    outs = prv.prover.ckt_outputs[:entry["n_copies"], :384].copy()
    outs[3, 17] = (outs[3, 17] + 1) % 256
    with open(tmp_path / "output.txt", "w") as fh:
        for row in outs.tolist():
            fh.write(" ".join(str(v) for v in row) + "\n")
    _, ver = gkr.VerifierClass.from_pws(parsed, entry["n_copies"])
    ver.vole_dir = str(tmp_path)
    ver.manifest_path = str(tmp_path / "orion_commit.json")
    ver.output_path = str(tmp_path / "output.txt")
    with pytest.raises(ValueError):
        ver.run(proof)


def test_edit_semantics_on_device(bridge, torch_cuda):
    """Circuit evaluation on the GPU implements the edits: gray keeps Y and
    sets chroma to 128, invert is 255 - x, crop multiplies by the keep bit,
    mask replicates samples 136 / 292 / 356."""
    n = 37
    for name in ("gray", "invert", "crop", "mask"):
        parsed = circuits.circuit(name)
        priv = random_private_inputs(name, parsed, n, seed=9)
        rows = full_input_rows(parsed, priv)
        _, layers = gkr.parse_layers(parsed)
        cp = gkr.CircuitProver(n, len(parsed[0]), layers)
        cp.set_inputs(rows.tolist())
        out = cp.ckt_outputs[:n, :384].astype(np.int64)
        if name == "gray":
            assert np.array_equal(out[:, :256], rows[:, :256]) and (out[:, 256:] == 128).all()
        elif name == "invert":
            assert np.array_equal(out, 255 - rows[:, :384])
        elif name == "crop":
            assert np.array_equal(out, rows[:, :384] * rows[:, 384:385])
        else:
            want = np.concatenate([np.repeat(rows[:, 136:137], 256, 1), np.repeat(rows[:, 292:293], 64, 1),
                                   np.repeat(rows[:, 356:357], 64, 1)], axis=1)
            assert np.array_equal(out, want)
        assert (cp.ckt_outputs[n:] == 0).all() and (cp.ckt_outputs[:, 384:] == 0).all()
        cp.close()
