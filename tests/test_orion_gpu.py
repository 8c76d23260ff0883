"""GPU parity of the signer's Orion commitment (tp_orion_commit,
tp_orion_encode): the roots and codewords equal the reference's
(tests/golden/orion_ref.json, from the oracle/_ref build of the reference's
layout, mask, expander code and Merkle tree) and the oracle's on fresh
inputs, seeds and columns."""
import hashlib
import json
import os

import numpy as np
import pytest

from truepix_amd import orion
from truepix_amd.fp2_link import commit_binding

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = (1 << 61) - 1


@pytest.fixture(scope="module")
def gold():
    with open(os.path.join(GOLD, "orion_ref.json")) as fh:
        return json.load(fh)


def inputs_for(copies, seed):
    return np.random.Generator(np.random.PCG64(seed)).integers(0, 256, size=(copies, 384),
                                                               dtype=np.uint8)


def test_commit_matches_reference_roots(bridge, gold):
    for case in gold["commit"]:
        print("orion commit copies", case["copies"], flush=True)
        m, k = orion.commit(inputs_for(case["copies"], case["input_seed"]), case["seed"])
        assert (m, k) == (case["masked_root"], case["mask_root"]), case["copies"]


def test_encode_matches_reference_codewords(bridge, gold):
    for case in gold["encode"]:
        n = case["n"]
        col = np.random.Generator(np.random.PCG64(n)).integers(0, P, size=(n, 2), dtype=np.uint64)
        cw = orion.encode(col[None])[0]
        assert cw.shape[0] == case["cw_len"], n
        assert hashlib.sha256(cw.tobytes()).hexdigest() == case["sha256"], n


def test_encode_many_columns_equals_oracle(bridge, oracle):
    """256 different columns in one launch set: each equals its own oracle
    codeword."""
    for n in (64, 1000):
        cols = np.random.Generator(np.random.PCG64(5 + n)).integers(0, P, size=(256, n, 2),
                                                                    dtype=np.uint64)
        got = orion.encode(cols)
        for c in (0, 1, 127, 128, 200, 255):
            assert np.array_equal(got[c], oracle.orion_encode(cols[c])), (n, c)


def test_commit_equals_oracle_on_new_inputs(bridge, oracle):
    for copies, seed in ((2, 1), (9, 2 ** 64 - 3), (100, 31337), (257, 8)):
        x = inputs_for(copies, copies)
        m, k, _ = oracle.orion_commit(x, seed)
        assert orion.commit(x, seed) == (m, k), copies


def test_manifest_binds_the_roots(bridge):
    x = inputs_for(37, 1)
    man = orion.manifest(x, 99)
    m, k = orion.commit(x, 99)
    assert (man["masked_root"], man["mask_root"]) == (m, k)
    assert man["n"] == 64 * 512 and man["padded_copies"] == 64
    assert commit_binding(man)[-2:] == [m, k]


@pytest.mark.parametrize("copies,seed,group,nq", [(1, 5, True, 8), (3, 7, False, 16), (16, 2024, True, 64),
                                                  (300, 12345, False, 200), (1350, 777, True, 1222),
                                                  (3600, 31, False, 300)])
def test_open_matches_oracle(bridge, oracle, copies, seed, group, nq):
    """The device opening (tp_orion_open) gives the oracle's OpeningProof
    bytes and answer; the oracle's bytes are pinned to the reference's own
    serializer (tests/test_orion_cpu.py)."""
    x = inputs_for(copies, copies + 3)
    _, _, cw = oracle.orion_commit(x, seed)
    pc = 1
    while pc < copies:
        pc <<= 1
    lg = (pc * 512).bit_length() - 1
    rng = np.random.Generator(np.random.PCG64(copies))
    r = rng.integers(0, (1 << 61) - 1, size=(lg, 2), dtype=np.uint64)
    q = rng.integers(0, cw, size=nq, dtype=np.int64)
    q[:3] = (0, cw - 1, 0)   # both ends, and a repeated position
    got, ans = orion.opening(x, seed, r, q, use_group1=group)
    want, want_ans = oracle.orion_open(x, seed, r, q, use_group1=group)
    assert ans == want_ans
    assert got == want
    assert orion.parse_opening(got)["answer"] == ans


def test_open_rejects_bad_points_and_queries(bridge, oracle):
    x = inputs_for(4, 1)
    _, _, cw = oracle.orion_commit(x, 3)
    r = np.ones((11, 2), np.uint64)
    with pytest.raises(RuntimeError):
        orion.opening(x, 3, r[:10], [0])          # lg n = 11 challenges needed
    with pytest.raises(RuntimeError):
        orion.opening(x, 3, r, [cw])              # past the codeword


def _root_bytes(hex_root):
    words = [int(hex_root[16 * i:16 * i + 16], 16) for i in range(4)]
    out = bytearray(32)
    for w, off in zip(words, (8, 0, 24, 16)):
        out[off:off + 8] = w.to_bytes(8, "little")
    return bytes(out)


@pytest.mark.parametrize("copies,seed,group", [(1, 5, True), (16, 2024, False), (1350, 777, True)])
def test_open_nizk_draws_the_reference_queries(bridge, oracle, copies, seed, group):
    """tp_orion_open_nizk draws the positions the reference's NIZK transcript
    draws (the oracle's draw is pinned to it) and opens them."""
    x = inputs_for(copies, copies + 5)
    m, k, cw = oracle.orion_commit(x, seed)
    pc = 1 << max(0, (copies - 1).bit_length())
    lg = (pc * 512).bit_length() - 1
    r = np.random.Generator(np.random.PCG64(copies + 1)).integers(0, (1 << 61) - 1, size=(lg, 2),
                                                                  dtype=np.uint64)
    q = oracle.orion_draw_queries(_root_bytes(m if group else k), pc * 512, group, cw, r)
    got, ans = orion.opening_nizk(x, seed, r, use_group1=group)
    want, want_ans = oracle.orion_open(x, seed, r, q, use_group1=group)
    assert orion.parse_opening(got)["col_q"] == q
    assert (got, ans) == (want, want_ans)


def _open_args(copies, point_seed, cw_len):
    """The point of the golden openings (tests/golden/make_golden_orion.py)."""
    pc = 1 << max(0, (copies - 1).bit_length())
    lg = (pc * 512).bit_length() - 1
    rng = np.random.Generator(np.random.PCG64(point_seed))
    return rng.integers(0, P, size=(lg, 2), dtype=np.uint64)


def test_open_nizk_virgo_matches_reference_transcripts(bridge, gold):
    """The Virgo code-switching transcript of each golden NIZK opening (every
    challenge, round polynomial, v_u, v_v, predicate and the verdict) hashes
    to the words the reference's own generate_circuit, zk_prover and
    zk_verifier checks produced (oracle/ref_harness_virgo.cpp)."""
    cw_of = {c["copies"]: c["cw_len"] for c in gold["commit"]}
    for case in gold["virgo"]:
        copies = case["copies"]
        r = _open_args(copies, case["point_seed"], cw_of[copies])
        print("virgo golden", copies, case["use_group1"], flush=True)
        proof, ans, words = orion.opening_nizk_virgo(inputs_for(copies, case["input_seed"]),
                                                     case["seed"], r, bool(case["use_group1"]))
        assert proof is not None and words[-1] == 1, copies
        assert words.size == case["words"] and [int(x) for x in words[1:1 + words[0]]] == case["bits"]
        assert hashlib.sha256(words.astype("<u8").tobytes()).hexdigest() == case["sha256"], copies
        plain, plain_ans = orion.opening_nizk(inputs_for(copies, case["input_seed"]), case["seed"], r,
                                              bool(case["use_group1"]))
        assert (proof, ans) == (plain, plain_ans)


@pytest.mark.parametrize("copies,seed,group", [(2, 9, False), (7, 31, True), (64, 5, False),
                                               (700, 2 ** 61 + 3, True), (3600, 11, False),
                                               (8040, 12, True), (20000, 13, False)])
def test_open_nizk_virgo_equals_oracle(bridge, oracle, copies, seed, group):
    """The code-switching transcript equals the oracle's (pinned to the
    reference) on fresh frames, both polynomials, up to a 20000-copy frame
    (a 2^17-entry message, 2^19-gate layers: eq half-tables of 2^10 in LDS)."""
    x = inputs_for(copies, copies + 17)
    pc = 1 << max(0, (copies - 1).bit_length())
    lg = (pc * 512).bit_length() - 1
    r = np.random.Generator(np.random.PCG64(copies + 3)).integers(0, P, size=(lg, 2), dtype=np.uint64)
    print("virgo oracle", copies, group, flush=True)
    proof, _, words = orion.opening_nizk_virgo(x, seed, r, group)
    assert proof is not None and words[-1] == 1
    o = orion.parse_opening(proof)
    want = oracle.orion_virgo(o["com_root"], pc * 512, group, r, o["combined_root"],
                              o["combined_message"])
    assert words.size == want.size
    assert np.array_equal(words, want)
    ms, launches, layers = orion.virgo_stats()
    assert launches == 2 and layers == int(words[0]) - 1 and ms > 0


def test_packings_match_reference(bridge, gold):
    """The invert, mask and crop circuits' packings on the device: roots, NIZK
    opening bytes and code-switching transcript equal the reference's
    (orion_ref.json); the gray packing is restored after."""
    try:
        for case in gold["layouts"]:
            lay = tuple(case["layout"])
            x = np.random.Generator(np.random.PCG64(case["input_seed"])).integers(
                0, 256, size=(case["copies"], lay[0]), dtype=np.uint8)
            print("packing", case["name"], lay, case["copies"], flush=True)
            assert orion.commit(x, case["seed"], layout=lay) == (case["masked_root"], case["mask_root"])
            r = _open_args(case["copies"], case["point_seed"], case["cw_len"])
            proof, _, words = orion.opening_nizk_virgo(x, case["seed"], r, bool(case["use_group1"]),
                                                       layout=lay)
            assert proof is not None and len(proof) == case["proof_len"]
            assert hashlib.sha256(proof).hexdigest() == case["proof_sha256"], case["name"]
            assert hashlib.sha256(words.astype("<u8").tobytes()).hexdigest() == case["virgo_sha256"]
            man = orion.manifest(x, case["seed"], layout=lay)
            assert (man["copy_size"], man["const_idx"], man["const_val"]) == lay
    finally:
        orion.commit(inputs_for(2, 1), 1)   # back to the gray packing
    with pytest.raises(ValueError, match="packing"):   # the constant inside the samples
        orion.commit(np.zeros((2, 385), np.uint8), 1, layout=(385, 384, 128))
