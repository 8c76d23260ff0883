"""GPU parity of the lockstep batch drivers (tp_pool_create_batched): frames
proved together, one launch per protocol step across the batch, give the
same proof bytes as the reference prover (golden fixtures from
TP/libTruePix/circuitnizkvec_vole.py via tests/golden/make_golden_reference.py)
and as the same frames proved alone by a stream worker; a failing frame does
not disturb the others; admission keeps the admitted table bytes within the
budget."""
import bz2
import hashlib
import json
import os

import numpy as np
import pytest

from truepix_amd import circuits, frames
from truepix_amd.fp2_link import (commit_binding, crop_bits, make_manifest, random_private_inputs,
                                  read_commit_manifest)

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = (1 << 61) - 1
NAMES = ("gray", "crop", "mask", "invert")


def golden():
    """The golden frames of the four edit circuits (the multi-layer cases are
    the Python-driven prover's)."""
    with open(os.path.join(GOLD, "golden_proofs.json")) as fh:
        idx = json.load(fh)
    return idx["pid"], [e for e in idx["cases"] if e["circuit"] in NAMES]


def check_gold(entry, res):
    assert len(res.proof) == entry["proof_len"], entry["name"]
    if "proof_file" in entry:
        with open(os.path.join(GOLD, entry["proof_file"]), "rb") as fh:
            assert res.proof == bz2.decompress(fh.read()), entry["name"]
    assert hashlib.sha256(res.proof).hexdigest() == entry["proof_sha256"], entry["name"]
    assert [[e.real, e.imag] for e in res.input_point] == entry["input_point"], entry["name"]
    assert (res.input_value.real, res.input_value.imag) == tuple(entry["input_value"]), entry["name"]


@pytest.fixture(scope="module")
def circs(bridge):
    c = {n: frames.FrameCircuit(circuits.circuit(n)) for n in NAMES}
    yield c
    for v in c.values():
        v.close()


def gold_specs(circs, reps=1):
    pid, entries = golden()
    vole = frames.vole_prover_arrays(os.path.join(GOLD, "vole_ext_prover_%d.json" % pid))
    out = []
    for rep in range(reps):
        for e in (entries if rep % 2 == 0 else list(reversed(entries))):
            parsed = circuits.circuit(e["circuit"])
            priv = random_private_inputs(e["circuit"], parsed, e["n_copies"], seed=e["input_seed"])
            binding = commit_binding(read_commit_manifest(os.path.join(GOLD, e["manifest"])))
            out.append((e, (circs[e["circuit"]], priv, pid, vole, binding)))
    return out


def synth_specs(circs, copies, seed):
    """Frames of every circuit at the given copy counts: uint8 samples from
    PCG64(seed), crop with its keep bit, VOLE U / V uniform nonzero."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k, n in enumerate(copies):
        name = NAMES[k % 4]
        c = circs[name]
        arr = rng.integers(0, 256, size=(n, c.n_private), dtype=np.uint8)
        if name == "crop":
            arr[:, -1] = crop_bits(n, 0.2, 0.8, 0.2, 0.8).astype(np.uint8)
        vole = tuple(rng.integers(1, P, size=500, dtype=np.int64).astype(np.uint64) for _ in range(4))
        out.append((c, arr, 30000 + k, vole, commit_binding(make_manifest(n, seed=k))))
    return out


def test_batch_matches_reference_golden_proofs(circs):
    """Every golden frame, twice, queued at once to one lockstep driver: the
    frames share launch sets and each proof is the reference's."""
    specs = gold_specs(circs, reps=2)
    with frames.FramePool(1, batched=True, max_batch=64, host_threads=8) as pool:
        ids = pool.submit_many([sp for _, sp in specs])
        print("batch submitted", len(ids), flush=True)
        res = [(e, pool.result(j)) for (e, _), j in zip(specs, ids)]
    for e, r in res:
        check_gold(e, r)
    assert all(r.batch == len(specs) for _, r in res)     # one launch set


def test_c4_batch_of_256_with_golden_frames_embedded(circs):
    """A C4-sized lockstep batch: 256 frames of the four circuits at 64x64,
    240p, 480p and 720p (16, 300, 1350, 3600 copies), with every golden frame
    (5 to 8040 copies) among them.  Each golden frame's proof is the
    reference's; every other frame's equals the same frame proved alone."""
    gold = gold_specs(circs)
    rng = np.random.Generator(np.random.PCG64(21))
    sizes = [int(x) for x in rng.choice([16, 300, 1350, 3600], size=256 - len(gold))]
    synth = synth_specs(circs, sizes, seed=22)
    order = rng.permutation(256)
    allspecs = [("gold", e, sp) for e, sp in gold] + [("synth", k, sp) for k, sp in enumerate(synth)]
    allspecs = [allspecs[i] for i in order]
    with frames.FramePool(4) as solo:
        ids = [solo.submit(*sp) for kind, _tag, sp in allspecs if kind == "synth"]
        want = iter([solo.result(j).proof for j in ids])
    print("solo done", flush=True)
    with frames.FramePool(1, batched=True, max_batch=256, host_threads=16) as pool:
        ids = pool.submit_many([sp for _, _, sp in allspecs])
        print("c4 batch submitted", len(ids), flush=True)
        got = [pool.result(j) for j in ids]
    n_gold = 0
    for (kind, tag, _), r in zip(allspecs, got):
        if kind == "gold":
            check_gold(tag, r)
            n_gold += 1
        else:
            assert r.proof == next(want), tag
    assert n_gold == len(gold) and max(r.batch for r in got) == 256


def test_batch_equals_stream_worker_on_mixed_shapes(circs):
    """Uneven frames (2 to 3600 copies, four circuits) in lockstep: each proof,
    input point and value equal the same frame proved alone."""
    copies = [2, 3, 16, 300, 1350, 3600, 5, 17, 64, 1000, 129, 2]
    specs = synth_specs(circs, copies, seed=11)
    with frames.FramePool(1) as solo:
        want = [solo.result(solo.submit(*sp)) for sp in specs]
    print("solo done", flush=True)
    with frames.FramePool(2, batched=True, max_batch=8, host_threads=8) as pool:
        jobs = [pool.submit(*sp) for sp in specs]
        got = [pool.result(j) for j in jobs]
        mem = pool.memory()
    for k, (a, b) in enumerate(zip(want, got)):
        assert a.proof == b.proof, (k, copies[k])
        assert a.input_point == b.input_point and a.input_value == b.input_value, k
    assert mem["frames"] == len(specs) and mem["in_use"] == 0
    assert all(1 <= r.batch <= 8 for r in got)


def test_batch_failing_frame_leaves_the_others(circs):
    specs = synth_specs(circs, [300, 64, 1350, 16], seed=12)
    bad = list(specs[1])
    bad[3] = tuple(a[:20] for a in bad[3])                 # too few VOLE correlations
    specs[1] = tuple(bad)
    with frames.FramePool(1) as solo:
        want = {k: solo.result(solo.submit(*sp)).proof for k, sp in enumerate(specs) if k != 1}
    with frames.FramePool(1, batched=True, max_batch=8, host_threads=4) as pool:
        jobs = [pool.submit(*sp) for sp in specs]
        with pytest.raises(RuntimeError, match="VOLE instance index out of range"):
            pool.result(jobs[1])
        for k in (0, 2, 3):
            assert pool.result(jobs[k]).proof == want[k], k
        # the driver keeps serving after a failed frame
        assert pool.result(pool.submit(*specs[0])).proof == want[0]


def test_batch_admission_respects_the_budget(circs):
    """A budget of one and a half large frames: batches stay within it, every frame
    still completes with its solo proof; a frame larger than the budget fails
    with a clear error instead of waiting forever."""
    specs = synth_specs(circs, [1350, 1350, 300, 1350, 16, 1350], seed=13)
    with frames.FramePool(1) as solo:
        want = [solo.result(solo.submit(*sp)) for sp in specs]
    fp = max(r.footprint for r in want)
    assert fp > 0
    budget = int(fp * 1.5)
    with frames.FramePool(2, batched=True, max_batch=64, budget_bytes=budget, host_threads=4) as pool:
        jobs = [pool.submit(*sp) for sp in specs]
        got = [pool.result(j) for j in jobs]
        mem = pool.memory()
        big = synth_specs(circs, [8040], seed=14)[0]
        jid = pool.submit(*big)
        with pytest.raises(RuntimeError, match="over the pool's budget"):
            pool.result(jid)
    for a, b in zip(want, got):
        assert a.proof == b.proof
    assert mem["budget"] == budget and 0 < mem["peak"] <= budget and mem["in_use"] == 0
    assert mem["batches"] >= 3


def test_batch_wide_outputs_take_the_u64_path(circs):
    """A circuit whose outputs exceed a byte (MUL / ADD / SUB / NOT / XOR over
    private inputs and a constant) batched beside byte-valued frames: its
    frames take the u64 outputs, the others the byte outputs, and every proof
    equals the frame proved alone."""
    inputs = [None] * 12 + [1000]
    ops = ("MUL", "ADD", "SUB", "NOT", "XOR", "NAND", "OR", "PASS")
    gates = [(ops[g % len(ops)], g % 13, (5 * g + 3) % 13, None) for g in range(40)]
    wide = frames.FrameCircuit([inputs, gates])
    rng = np.random.Generator(np.random.PCG64(15))
    specs = []
    for k, n in enumerate([40, 3, 300, 17]):
        arr = rng.integers(0, 256, size=(n, wide.n_private), dtype=np.uint8)
        vole = tuple(rng.integers(1, P, size=500, dtype=np.int64).astype(np.uint64) for _ in range(4))
        specs.append((wide, arr, 40000 + k, vole, commit_binding(make_manifest(n, seed=k))))
    specs = [specs[0]] + synth_specs(circs, [64, 300], seed=16) + specs[1:]
    with frames.FramePool(1) as solo:
        want = [solo.result(solo.submit(*sp, keep_outputs=True)) for sp in specs]
    assert int(want[0].outputs.max()) > 255
    with frames.FramePool(1, batched=True, max_batch=16, host_threads=4) as pool:
        jobs = [pool.submit(*sp, keep_outputs=True) for sp in specs]
        got = [pool.result(j) for j in jobs]
    for k, (a, b) in enumerate(zip(want, got)):
        assert a.proof == b.proof, k
        assert np.array_equal(a.outputs, b.outputs), k
    wide.close()


def test_schedule_tuner_settles_and_keeps_proofs(circs):
    """The lockstep schedule's autotuner tries three early-claim sizes, then
    three fold sizes, on real batches and settles; every schedule it tries,
    and the extremes fixed by hand, give the frames' solo proofs."""
    specs = synth_specs(circs, [1350] * 6 + [300, 16], seed=17)
    with frames.FramePool(1) as solo:
        want = [solo.result(solo.submit(*sp)).proof for sp in specs]
    with frames.FramePool(1, batched=True, max_batch=8, host_threads=8) as pool:
        for rep in range(7):
            ids = pool.submit_many(specs)
            assert [pool.result(j).proof for j in ids] == want, rep
        t = pool.tuning()
    print("tuning", t, flush=True)
    assert t["phase"] == 2
    assert all(c > 0 for c in t["claim_ns_per_term"]) and all(c > 0 for c in t["fold_ns_per_byte"])
    assert t["claim_items"] in (1 << 14, 1 << 16, 1 << 18) and t["fold_cap"] in (256, 1024, 4096)
    with frames.FramePool(1, batched=True, max_batch=8, host_threads=8) as pool:
        for claim, fold in ((1024, 16), (1 << 18, 65536)):
            pool.fix_schedule(claim, fold)
            ids = pool.submit_many(specs)
            assert [pool.result(j).proof for j in ids] == want, (claim, fold)
        assert pool.tuning()["claim_items"] == 1 << 18
        with pytest.raises(RuntimeError):
            pool.fix_schedule(1 << 20, 1024)


def test_sixteen_stream_workers_grow_buffers_while_others_answer_resident_rounds(circs):
    """Sixteen stream workers take frames of rising size, so workers grow
    their buffers (frees that synchronize the device) while other workers'
    resident late-round kernels wait on their host threads for challenges.
    Every frame completes and equals the same frame proved alone."""
    copies = [16] * 16 + [300] * 16 + [1350] * 24 + [3600] * 24 + [16, 3600] * 8
    specs = synth_specs(circs, copies, seed=31)
    with frames.FramePool(1) as solo:
        want = [solo.result(solo.submit(*sp)).proof for sp in specs]
    print("solo done", flush=True)
    with frames.FramePool(16) as pool:
        jobs = [pool.submit(*sp) for sp in specs]
        got = [pool.result(j).proof for j in jobs]
    for k, (a, b) in enumerate(zip(want, got)):
        assert a == b, (k, copies[k])
