"""GPU parity of the whole-frame prover (tp_pool_*): for fixed inputs, VOLE
correlations, manifest and pid, the proof bytes equal the reference prover's
(golden fixtures from TP/libTruePix/circuitnizkvec_vole.py, run by
tests/golden/make_golden_reference.py), alone and inside a mix of frames
proved concurrently by several workers."""
import bz2
import hashlib
import json
import os
import shutil

import numpy as np
import pytest

from truepix_amd import circuits, frames, gkr
from truepix_amd.extfield import Fp2
from truepix_amd.fp2_link import (commit_binding, full_input_rows, random_private_inputs,
                                  read_commit_manifest)

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cases():
    """The golden frames of the four edit circuits."""
    with open(os.path.join(GOLD, "golden_proofs.json")) as fh:
        idx = json.load(fh)
    return idx["pid"], [e for e in idx["cases"] if e["circuit"] in circuits.CIRCUITS]


def gold_inputs(entry):
    parsed = circuits.circuit(entry["circuit"])
    priv = random_private_inputs(entry["circuit"], parsed, entry["n_copies"], seed=entry["input_seed"])
    return parsed, priv


def check_against_gold(entry, proof, point, value):
    assert len(proof) == entry["proof_len"], entry["name"]
    if "proof_file" in entry:
        with open(os.path.join(GOLD, entry["proof_file"]), "rb") as fh:
            assert proof == bz2.decompress(fh.read()), entry["name"]
    assert hashlib.sha256(proof).hexdigest() == entry["proof_sha256"], entry["name"]
    assert [[e.real, e.imag] for e in point] == entry["input_point"], entry["name"]
    assert (value.real, value.imag) == tuple(entry["input_value"]), entry["name"]


@pytest.mark.parametrize("name", ["gray16", "inv8", "crop12", "mask5", "gray1350", "gray3600",
                                  "inv3600", "crop8040", "mask8040"])
def test_frame_prover_matches_reference(bridge, tmp_path, name):
    pid, entries = cases()
    entry = next(e for e in entries if e["name"] == name)
    shutil.copy(os.path.join(GOLD, "vole_ext_prover_%d.json" % pid), tmp_path)
    parsed, priv = gold_inputs(entry)
    _, prv = frames.FrameProverNIZK.from_pws(parsed, entry["n_copies"])
    prv.pid = pid
    prv.vole_dir = str(tmp_path)
    prv.manifest_path = os.path.join(GOLD, entry["manifest"])
    print("frame case", name, flush=True)
    proof = prv.run(full_input_rows(parsed, priv).tolist())
    check_against_gold(entry, proof, prv.final_input_point, prv.final_input_value)
    # the device outputs are the edited frame (the verifier's output.txt)
    out = prv.ckt_outputs[:entry["n_copies"], :384].astype(np.int64)
    rows = full_input_rows(parsed, priv)
    if entry["circuit"] == "invert":
        assert np.array_equal(out, 255 - rows[:, :384])
    elif entry["circuit"] == "gray":
        assert np.array_equal(out[:, :256], rows[:, :256]) and (out[:, 256:] == 128).all()
    prv.circuit.close()


def test_frames_in_a_mix_equal_their_solo_proofs(bridge):
    """Alone-vs-mix parity: every golden frame, submitted twice and interleaved
    to a pool of three workers, yields the reference's proof bytes."""
    pid, entries = cases()
    vole = frames.vole_prover_arrays(os.path.join(GOLD, "vole_ext_prover_%d.json" % pid))
    circs = {}
    jobs = []
    with frames.FramePool(3) as pool:
        for rep in range(2):
            for entry in (entries if rep == 0 else list(reversed(entries))):
                name = entry["circuit"]
                if name not in circs:
                    circs[name] = frames.FrameCircuit(circuits.circuit(name))
                _, priv = gold_inputs(entry)
                binding = commit_binding(read_commit_manifest(os.path.join(GOLD, entry["manifest"])))
                jobs.append((entry, pool.submit(circs[name], priv, pid, vole, binding)))
        print("mix submitted", len(jobs), flush=True)
        for entry, jid in jobs:
            res = pool.result(jid)
            check_against_gold(entry, res.proof, res.input_point, res.input_value)
            assert res.t_submit <= res.t_start <= res.t_end
    for c in circs.values():
        c.close()


def test_pool_keeps_a_dropped_circuit_alive(bridge):
    """A frame queued with a circuit object the caller drops at once (and a
    circuit closed while its frames are queued) still proves with the
    reference's bytes: the pool and the library hold the circuit until the
    job is released."""
    import gc
    pid, entries = cases()
    vole = frames.vole_prover_arrays(os.path.join(GOLD, "vole_ext_prover_%d.json" % pid))
    picked = [e for e in entries if e["name"] in ("gray1350", "crop12", "inv8")]
    with frames.FramePool(1) as pool:
        jobs = []
        for entry in picked:
            _, priv = gold_inputs(entry)
            binding = commit_binding(read_commit_manifest(os.path.join(GOLD, entry["manifest"])))
            c = frames.FrameCircuit(circuits.circuit(entry["circuit"]))
            jobs.append((entry, pool.submit(c, priv, pid, vole, binding)))
            c.close()                # closed by the caller while the frame is queued
            del c
        gc.collect()
        for entry, jid in jobs:
            res = pool.result(jid)
            check_against_gold(entry, res.proof, res.input_point, res.input_value)
        _, priv = gold_inputs(picked[0])
        binding = commit_binding(read_commit_manifest(os.path.join(GOLD, picked[0]["manifest"])))
        res = pool.result(pool.submit(frames.FrameCircuit(circuits.circuit(picked[0]["circuit"])),
                                      priv, pid, vole, binding))
        check_against_gold(picked[0], res.proof, res.input_point, res.input_value)


def test_frame_errors_are_reported(bridge):
    pid, entries = cases()
    entry = next(e for e in entries if e["name"] == "inv8")
    vole = frames.vole_prover_arrays(os.path.join(GOLD, "vole_ext_prover_%d.json" % pid))
    binding = commit_binding(read_commit_manifest(os.path.join(GOLD, entry["manifest"])))
    circ = frames.FrameCircuit(circuits.circuit("invert"))
    _, priv = gold_inputs(entry)
    with frames.FramePool(1) as pool:
        with pytest.raises(ValueError):
            pool.submit(circ, priv[:, :10], pid, vole, binding)
        with pytest.raises(RuntimeError):
            pool.submit(circ, priv[:1], pid, vole, binding)           # one copy
        short = tuple(a[:20] for a in vole)                           # too few correlations
        jid = pool.submit(circ, priv, pid, short, binding)
        with pytest.raises(RuntimeError, match="VOLE instance index out of range"):
            pool.result(jid)
        # the worker keeps serving after a failed frame
        res = pool.result(pool.submit(circ, priv, pid, vole, binding))
        check_against_gold(entry, res.proof, res.input_point, res.input_value)
    circ.close()


def test_python_and_frame_provers_agree_on_new_inputs(bridge, tmp_path):
    """Outside the fixtures: the device frame prover and the Python-driven
    prover (gkr.CircuitProverNIZK) give the same proof for fresh inputs."""
    pid = 777
    from truepix_amd.fp2_link import make_manifest, write_manifest
    from truepix_amd.vole import ExtVOLEGenerator
    gen = ExtVOLEGenerator(seed=5)
    gen.generate_deltas()
    gen.generate_vole_instances(500)
    gen.save_to_files(str(tmp_path / ("vole_ext_prover_%d.json" % pid)),
                      str(tmp_path / ("vole_ext_verifier_%d.json" % pid)))
    for name, n in (("crop", 37), ("mask", 100)):
        parsed = circuits.circuit(name)
        priv = random_private_inputs(name, parsed, n, seed=n)
        write_manifest(str(tmp_path / "m.json"), make_manifest(n, seed=n))
        rows = full_input_rows(parsed, priv).tolist()
        _, a = gkr.ProverClass.from_pws(parsed, n)
        _, b = frames.FrameProverNIZK.from_pws(parsed, n)
        for prv in (a, b):
            prv.pid = pid
            prv.vole_dir = str(tmp_path)
            prv.manifest_path = str(tmp_path / "m.json")
        pa = a.run(rows)
        pb = b.run(rows)
        assert pa == pb, name
        assert a.final_input_value == b.final_input_value
        b.circuit.close()
