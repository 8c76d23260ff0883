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

from oracle.oracle import Oracle, RefOrion

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
