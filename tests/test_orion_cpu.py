"""CPU checks of the Orion commitment oracle (oracle/fp2_oracle.c
or_orion_commit, or_orion_encode, or_sha256) against the reference's own
commitment pieces (tests/golden/orion_ref.json, written from the
oracle/_ref build of OZ/include/truepix/truepix_gkr.h, linear_code/ and
merkle_tree.cpp), against the live reference build where it exists, and
against what the mathematics fixes: SHA-256 known answers, the code's
linearity over Fp2 and its systematic prefix."""
import hashlib
import json
import os
import random

import numpy as np
import pytest

from oracle.oracle import Oracle, RefOrion, RefVirgo

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = (1 << 61) - 1


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.fixture(scope="module")
def gold():
    with open(os.path.join(GOLD, "orion_ref.json")) as fh:
        return json.load(fh)


def inputs_for(copies, seed):
    return np.random.Generator(np.random.PCG64(seed)).integers(0, 256, size=(copies, 384),
                                                               dtype=np.uint8)


def column_for(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.integers(0, P, size=(n, 2), dtype=np.uint64)


def test_sha256_known_answers(orc):
    rng = random.Random(9)
    for n in (0, 3, 55, 56, 63, 64, 65, 127, 128, 1000):
        data = bytes(rng.getrandbits(8) for _ in range(n))
        assert orc.sha256(data) == hashlib.sha256(data).digest(), n
    assert orc.sha256(b"abc").hex() == \
        "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"


def test_encode_matches_reference_codewords(orc, gold):
    for case in gold["encode"]:
        cw = orc.orion_encode(column_for(case["n"], case["n"]))
        assert cw.shape[0] == case["cw_len"], case["n"]
        assert [int(x) for x in cw[:4].reshape(-1)] == case["head"], case["n"]
        assert hashlib.sha256(cw.tobytes()).hexdigest() == case["sha256"], case["n"]


def _fp2_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def test_encode_is_linear_and_systematic(orc):
    """encode(a x + b y) = a encode(x) + b encode(y) over Fp2, and the
    codeword starts with the message (a dropped edge, a wrong weight part or
    a transposed weight breaks the first; a misplaced block the second)."""
    for n in (14, 40, 300):
        x, y = column_for(n, 100 + n), column_for(n, 200 + n)
        a, b = (12345, 678), (P - 5, 99)
        comb = np.zeros_like(x)
        for i in range(n):
            ax = _fp2_mul(a, (int(x[i, 0]), int(x[i, 1])))
            by = _fp2_mul(b, (int(y[i, 0]), int(y[i, 1])))
            comb[i] = ((ax[0] + by[0]) % P, (ax[1] + by[1]) % P)
        ex, ey, ec = orc.orion_encode(x), orc.orion_encode(y), orc.orion_encode(comb)
        assert np.array_equal(ex[:n], x)
        for k in range(ec.shape[0]):
            ax = _fp2_mul(a, (int(ex[k, 0]), int(ex[k, 1])))
            by = _fp2_mul(b, (int(ey[k, 0]), int(ey[k, 1])))
            assert (int(ec[k, 0]), int(ec[k, 1])) == ((ax[0] + by[0]) % P, (ax[1] + by[1]) % P), (n, k)


def test_commit_matches_reference_roots(orc, gold):
    for case in gold["commit"]:
        m, k, cw = orc.orion_commit(inputs_for(case["copies"], case["input_seed"]), case["seed"])
        assert (m, k, cw) == (case["masked_root"], case["mask_root"], case["cw_len"]), case["copies"]


def test_commit_depends_on_every_part(orc):
    """One changed sample changes the masked root only; another seed changes
    both."""
    x = inputs_for(16, 3)
    m0, k0, _ = orc.orion_commit(x, 5)
    y = x.copy()
    y[7, 100] ^= 1
    m1, k1, _ = orc.orion_commit(y, 5)
    assert m1 != m0 and k1 == k0
    m2, k2, _ = orc.orion_commit(x, 6)
    assert m2 != m0 and k2 != k0


@pytest.mark.skipif(not RefOrion.available(), reason="reference build absent (make -C oracle ref)")
def test_commit_matches_live_reference(orc):
    ref = RefOrion()
    for copies, seed in ((2, 1), (7, 2 ** 63 + 5), (64, 31337)):
        x = inputs_for(copies, seed % 1000)
        assert orc.orion_commit(x, seed) == ref.commit(x, seed), copies
    for n in (17, 250):
        col = column_for(n, 7 * n)
        assert np.array_equal(orc.orion_encode(col), ref.encode(col)), n


# ------------------------------------------------------------------ opening

def open_args(copies, point_seed, nq, cw_len):
    pc = 1
    while pc < copies:
        pc <<= 1
    lg = (pc * 512).bit_length() - 1
    rng = np.random.Generator(np.random.PCG64(point_seed))
    r = rng.integers(0, P, size=(lg, 2), dtype=np.uint64)
    q = rng.integers(0, cw_len, size=nq, dtype=np.int64)
    q[:2] = (0, cw_len - 1)
    return r, q


def test_open_matches_reference_proofs(orc, gold):
    cw_of = {c["copies"]: c["cw_len"] for c in gold["commit"]}
    for case in gold["open"]:
        r, q = open_args(case["copies"], case["point_seed"], case["queries"], cw_of[case["copies"]])
        proof, _ = orc.orion_open(inputs_for(case["copies"], case["input_seed"]), case["seed"], r, q,
                                  use_group1=bool(case["use_group1"]))
        assert len(proof) == case["proof_len"], case
        assert hashlib.sha256(proof).hexdigest() == case["proof_sha256"], case


def _eq_msb(r, k):
    """eq weights over k bits, r[0] on the top bit (Python big ints)."""
    out = [(1, 0)]
    for d in range(k):
        x = (int(r[d][0]), int(r[d][1]))
        nx = ((1 - x[0]) % P, (-x[1]) % P)
        # each new challenge takes the lowest bit, pushing r[0] up to the top
        out = [w for v in out for w in (_fp2_mul(v, nx), _fp2_mul(v, x))]
    return out


def _walk(leaf, pos, leaves, sib):
    node, v = leaf, pos + leaves
    while v != 1:
        s = sib[v ^ 1]
        node = hashlib.sha256((node + s) if v % 2 == 0 else (s + node)).digest()
        v //= 2
    return node


def _fe(x):
    return int(x[1]).to_bytes(8, "little") + int(x[0]).to_bytes(8, "little")   # (imag, real)


def test_open_proof_checks_out(orc):
    """What the reference verifier checks of an opening, with independent
    code: the combined codeword is the code of the combined message (the code
    is linear and systematic), each queried column's r0-combination is the
    combined codeword there, and both Merkle paths lead to the roots; the
    opened root is the commitment's."""
    from truepix_amd.orion import parse_opening
    for copies, seed, group in ((2, 3, True), (9, 41, False)):
        x = inputs_for(copies, copies)
        _, _, cw = orc.orion_commit(x, seed)
        r, q = open_args(copies, 100 + copies, 12, cw)
        proof, ans = orc.orion_open(x, seed, r, q, use_group1=group)
        o = parse_opening(proof)
        assert o["answer"] == ans and o["cw_len"] == cw and o["col_q"] == list(q)
        cm, cc = o["combined_message"], o["combined_codeword"]
        assert np.array_equal(orc.orion_encode(cm), cc)
        r0 = _eq_msb(r, 7)
        rows = 2 * o["msg_len"]
        for i, pos in enumerate(q):
            col = o["queried_cols"][i]
            s = (0, 0)
            for j in range(128):
                t = _fp2_mul(r0[j], (int(col[j][0]), int(col[j][1])))
                s = ((s[0] + t[0]) % P, (s[1] + t[1]) % P)
            assert s == (int(cc[pos][0]), int(cc[pos][1])), (copies, i)
            h = bytes(32)
            for j in range(64):
                h = hashlib.sha256(h + _fe(col[2 * j]) + _fe(col[2 * j + 1])).digest()
            assert _walk(h, int(pos), rows, o["com_siblings"]) == o["com_root"]
            leaf = hashlib.sha256(_fe(cc[pos]) + bytes(48)).digest()
            assert _walk(leaf, int(pos), rows, o["combined_siblings"]) == o["combined_root"]
        # the opened tree is the committed one: its root prints as the manifest's
        m, k, _ = orc.orion_commit(x, seed)
        words = [int.from_bytes(o["com_root"][a:a + 8], "little") for a in (8, 0, 24, 16)]
        assert "".join("%016x" % w for w in words) == (m if group else k)


@pytest.mark.skipif(not RefOrion.available(), reason="reference build absent (make -C oracle ref)")
def test_open_matches_live_reference(orc):
    ref = RefOrion()
    for copies, seed, group in ((4, 17, True), (40, 2 ** 62 + 9, False)):
        x = inputs_for(copies, seed % 1000)
        _, _, cw = orc.orion_commit(x, seed)
        r, q = open_args(copies, seed % 997, 33, cw)
        assert orc.orion_open(x, seed, r, q, use_group1=group)[0] == ref.open(x, seed, r, q, group)


def root_bytes(hex_root):
    """A digest's bytes from its root_hex form (u64 words at byte offsets 8, 0,
    24, 16)."""
    words = [int(hex_root[16 * i:16 * i + 16], 16) for i in range(4)]
    out = bytearray(32)
    for w, off in zip(words, (8, 0, 24, 16)):
        out[off:off + 8] = w.to_bytes(8, "little")
    return bytes(out)


def test_query_draw_matches_reference(orc, gold):
    """The opening's NIZK query draw (orion_fs.cpp's transcript) against the
    positions the reference's own transcript drew."""
    cw_of = {c["copies"]: c["cw_len"] for c in gold["commit"]}
    for case in gold["draw"]:
        copies = case["copies"]
        r, _ = open_args(copies, case["point_seed"], 2, cw_of[copies])
        m, k, cw = orc.orion_commit(inputs_for(copies, case["input_seed"]), case["seed"])
        pc = 1 << max(0, (copies - 1).bit_length())
        q = orc.orion_draw_queries(root_bytes(m if case["use_group1"] else k), pc * 512,
                                   bool(case["use_group1"]), cw, r)
        assert len(q) == case["queries"] == 1222 and q[:8] == case["head"], copies
        assert hashlib.sha256(np.asarray(q, "<i8").tobytes()).hexdigest() == case["sha256"], copies
        assert all(x % 2 == (0 if case["use_group1"] else 1) and 0 <= x < cw for x in q)


@pytest.mark.skipif(not RefOrion.available(), reason="reference build absent (make -C oracle ref)")
def test_query_draw_matches_live_reference(orc):
    ref = RefOrion()
    for copies, seed, group in ((2, 8, True), (77, 5, False)):
        x = inputs_for(copies, 3)
        m, k, cw = orc.orion_commit(x, seed)
        pc = 1 << max(0, (copies - 1).bit_length())
        r, _ = open_args(copies, copies, 2, cw)
        root = root_bytes(m if group else k)
        assert orc.orion_draw_queries(root, pc * 512, group, cw, r) == \
            ref.draw_queries(root, pc * 512, group, cw, r)


# ------------------------------------------------- code-switching proof

def _virgo_case(orc, copies, seed, iseed, group, pseed):
    """The oracle's NIZK opening of a seeded frame and its code-switching
    transcript."""
    from truepix_amd.orion import parse_opening
    x = inputs_for(copies, iseed)
    m, k, cw = orc.orion_commit(x, seed)
    r, _ = open_args(copies, pseed, 2, cw)
    pc = 1 << max(0, (copies - 1).bit_length())
    q = orc.orion_draw_queries(root_bytes(m if group else k), pc * 512, group, cw, r)
    proof, _ = orc.orion_open(x, seed, r, np.asarray(q, np.int64), use_group1=group)
    o = parse_opening(proof)
    args = (o["com_root"], pc * 512, group, r, o["combined_root"], o["combined_message"])
    return orc.orion_virgo(*args), args


def test_virgo_matches_reference_transcripts(orc, gold):
    """The oracle's Virgo code-switching transcript (every challenge, round
    polynomial, claim, predicate and the verdict) against the reference's
    (generate_circuit, zk_prover stepped in zk_verifier::verify's order, the
    verifier's predicates; tests/golden/orion_ref.json)."""
    for case in gold["virgo"]:
        words, _ = _virgo_case(orc, case["copies"], case["seed"], case["input_seed"],
                               bool(case["use_group1"]), case["point_seed"])
        assert words.size == case["words"] and words[-1] == case["verdict"] == 1, case["copies"]
        assert [int(x) for x in words[1:1 + words[0]]] == case["bits"]
        assert hashlib.sha256(words.astype("<u8").tobytes()).hexdigest() == case["sha256"], case


@pytest.mark.skipif(not RefVirgo.available(), reason="reference build absent (make -C oracle ref)")
def test_virgo_matches_live_reference(orc):
    ref = RefVirgo()
    for copies, seed, group in ((2, 4, False), (5, 8, True), (40, 2 ** 62 + 1, False), (520, 3, True)):
        words, args = _virgo_case(orc, copies, seed, copies, group, copies + 1)
        assert np.array_equal(words, ref.transcript(*args)), copies


def test_virgo_transcript_shape_and_closed_forms(orc):
    """What the algebra fixes about the transcript: the circuit of
    generate_circuit (layer 1 as wide as the message, the output layer 2^11
    for 1222 sorted positions inside the codeword), r_v zero at layer 1, and
    for relays and linear combinations a phase 2 of (0, -S_j, S_j) with
    S_{j+1} = S_j (1 - r_v[j]), S = 0 on the direct relays of layer 1."""
    words, args = _virgo_case(orc, 16, 2024, 4, True, 23)
    w = [int(x) for x in words]
    D = w[0]
    bits = w[1:1 + D]
    msg_len = args[1] // 128
    assert bits[0] == bits[1] == msg_len.bit_length() - 1 and bits[-1] == 11
    pos = 1 + D
    nq = w[pos]
    q = w[pos + 1:pos + 1 + nq]
    assert nq == 1222 and q == sorted(q)
    pos += 1 + nq + 4 * bits[-1] + 2
    for i in range(D - 1, 0, -1):
        b = bits[i - 1]
        ru = [tuple(w[pos + 2 * k:pos + 2 * k + 2]) for k in range(b)]
        rv = [tuple(w[pos + 2 * b + 2 * k:pos + 2 * b + 2 * k + 2]) for k in range(b)]
        pos += 4 * b
        if i == 1:
            assert all(x == (0, 0) for x in rv)
        pos += 6 * b   # phase 1
        p2 = [w[pos + 6 * k:pos + 6 * k + 6] for k in range(b)]
        pos += 6 * b
        s = (p2[0][4], p2[0][5])
        for k in range(b):
            a, bb, c = (p2[k][0], p2[k][1]), (p2[k][2], p2[k][3]), (p2[k][4], p2[k][5])
            assert a == (0, 0) and c == s and bb == ((-s[0]) % P, (-s[1]) % P), (i, k)
            s = _fp2_mul(s, ((1 - rv[k][0]) % P, (-rv[k][1]) % P))
        if i == 1:
            assert (p2[0][4], p2[0][5]) == (0, 0)
        pos += 4 + 6 + 4   # v_u v_v, three predicates, alpha beta
    assert pos + 5 == len(w) and w[-1] == 1


def test_virgo_checks_catch_planted_mistakes(orc):
    """The verifier's checks are live: a mistake planted in a round
    polynomial, in v_v, in a predicate or in the input layer's value makes
    the verdict 0."""
    try:
        for k in (1, 2, 3, 4):
            orc.orion_virgo_plant(k)
            words, _ = _virgo_case(orc, 16, 2024, 4, True, 23)
            assert words[-1] == 0, k
    finally:
        orc.orion_virgo_plant(0)
    words, _ = _virgo_case(orc, 16, 2024, 4, True, 23)
    assert words[-1] == 1


# ------------------------------------------------ the circuits' packings

def _layout_inputs(copies, seed, copy_size):
    return np.random.Generator(np.random.PCG64(seed)).integers(0, 256, size=(copies, copy_size),
                                                               dtype=np.uint8)


def test_packings_match_reference(orc, gold):
    """The invert, mask and crop circuits' packings (constant 255, no
    constant, 385 samples): commitment roots, NIZK opening bytes and the
    code-switching transcript against the reference's (orion_ref.json)."""
    from truepix_amd.orion import parse_opening
    try:
        for case in gold["layouts"]:
            orc.orion_set_layout(*case["layout"])
            x = _layout_inputs(case["copies"], case["input_seed"], case["layout"][0])
            m, k, cw = orc.orion_commit(x, case["seed"])
            assert (m, k, cw) == (case["masked_root"], case["mask_root"], case["cw_len"]), case["name"]
            r, _ = open_args(case["copies"], case["point_seed"], 2, cw)
            group = bool(case["use_group1"])
            pc = 1 << max(0, (case["copies"] - 1).bit_length())
            q = orc.orion_draw_queries(root_bytes(m if group else k), pc * 512, group, cw, r)
            proof, _ = orc.orion_open(x, case["seed"], r, np.asarray(q, np.int64), use_group1=group)
            assert hashlib.sha256(proof).hexdigest() == case["proof_sha256"], case["name"]
            o = parse_opening(proof)
            words = orc.orion_virgo(o["com_root"], pc * 512, group, r, o["combined_root"],
                                    o["combined_message"])
            assert hashlib.sha256(words.astype("<u8").tobytes()).hexdigest() == case["virgo_sha256"]
    finally:
        orc.orion_set_layout(384, 384, 128)


def test_packing_changes_only_what_it_should(orc):
    """A constant changes only the masked root (the mask covers samples
    only); dropping the constant changes it again; the crop packing's 385th
    sample enters the mask stream, so both roots change."""
    x = _layout_inputs(9, 3, 385)
    try:
        orc.orion_set_layout(384, 384, 128)
        g = orc.orion_commit(x[:, :384], 7)
        orc.orion_set_layout(384, 384, 255)
        i = orc.orion_commit(x[:, :384], 7)
        orc.orion_set_layout(384, -1, 0)
        n = orc.orion_commit(x[:, :384], 7)
        orc.orion_set_layout(385, -1, 0)
        c = orc.orion_commit(x, 7)
        assert g[1] == i[1] == n[1] and len({g[0], i[0], n[0]}) == 3
        assert c[0] not in (g[0], n[0]) and c[1] != g[1]
        with pytest.raises(ValueError):
            orc.orion_set_layout(385, 384, 128)   # the constant inside the samples
    finally:
        orc.orion_set_layout(384, 384, 128)


@pytest.mark.skipif(not RefOrion.available(), reason="reference build absent (make -C oracle ref)")
def test_packings_match_live_reference(orc):
    ref = RefOrion()
    try:
        for lay, copies, seed in (((384, 384, 255), 5, 2), ((384, -1, 0), 33, 4), ((385, -1, 0), 70, 6)):
            orc.orion_set_layout(*lay)
            RefOrion.set_layout(*lay)
            x = _layout_inputs(copies, seed, lay[0])
            assert orc.orion_commit(x, seed) == ref.commit(x, seed), lay
            _, _, cw = ref.commit(x, seed)
            r, q = open_args(copies, seed, 20, cw)
            assert orc.orion_open(x, seed, r, q, use_group1=True)[0] == ref.open(x, seed, r, q, True)
    finally:
        orc.orion_set_layout(384, 384, 128)
        RefOrion.set_layout(384, 384, 128)
