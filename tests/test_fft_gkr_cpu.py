"""CPU checks of the fft_gkr oracle (oracle/fp2_oracle.c or_fft_gkr, the
restatement of OZ/src/VPD/fft_circuit_GKR.cpp): its rand() stream equals
glibc's, and for every size it matches the golden values the reference's own
fft_gkr produced (tests/golden/make_golden_fftgkr.py) and, where the reference
builds here, the reference run side by side."""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

from oracle.oracle import RefFftGkr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fft_gkr_ref.json")
P = (1 << 61) - 1


def gold_cases():
    with open(GOLD) as fh:
        return json.load(fh)["cases"]


def layer_digests(layers, lg):
    n = 1 << lg
    raw = layers.tobytes()
    return {"ifft_out": hashlib.sha256(raw[:16 * n]).hexdigest(),
            "scaled": hashlib.sha256(raw[16 * n:32 * n]).hexdigest(),
            "sums": hashlib.sha256(raw[32 * n:32 * n + 64 * 16]).hexdigest()}


def test_glibc_rand_restatement(oracle):
    libc = ctypes.CDLL(None)
    for seed in (1, 2, 12345):
        libc.srand(seed)
        assert oracle.glibc_rand(seed, 600) == [libc.rand() for _ in range(600)]


@pytest.mark.parametrize("case", gold_cases(), ids=lambda c: "lg%d" % c["lg_size"])
def test_oracle_matches_reference_golden(oracle, case):
    lg = case["lg_size"]
    r = oracle.fft_gkr(lg)
    assert r["ok"] == case["ok"] == 1
    assert r["proof_size"] == case["proof_size"]
    assert [int(x) for x in r["claims"]] == case["claims"]
    assert layer_digests(r["layers"], lg) == case["layer_sha256"]
    assert r["n_polys"] == (lg + 6) + lg + 2 * lg * lg


@pytest.mark.skipif(not RefFftGkr.available(), reason="reference harness not built here")
def test_oracle_matches_reference_run(oracle):
    ref = RefFftGkr()
    for lg in (3, 8, 13):
        a, b = oracle.fft_gkr(lg, with_polys=False), ref.run(lg)
        assert a["ok"] == b["ok"] == 1 and a["proof_size"] == b["proof_size"]
        assert np.array_equal(a["claims"], b["claims"])
        assert np.array_equal(a["layers"], b["layers"])


def test_round_polynomials_chain(oracle):
    """Each round polynomial's s(0) + s(1) equals the previous claim: the
    first round of the addition layer against the claim from the 64 sums."""
    lg = 5
    r = oracle.fft_gkr(lg)
    polys = r["polys"].reshape(-1, 6)
    a = (int(polys[0, 0]), int(polys[0, 1]))
    b = (int(polys[0, 2]), int(polys[0, 3]))
    c = (int(polys[0, 4]), int(polys[0, 5]))
    s01 = ((2 * c[0] + a[0] + b[0]) % P, (2 * c[1] + a[1] + b[1]) % P)
    assert s01 == (int(r["claims"][0]), int(r["claims"][1]))
