"""CPU checks of the frame prover's host pieces inside libtruepix_hip.so: the
transcript hash (SHA-256, SHA-NI path when the CPU has it) against hashlib,
and the proof writer against CPython's pickle (the reference's proof bytes,
TP/libTruePix/fiatshamir.py:234-237)."""
import bz2
import ctypes
import hashlib
import json
import os
import pickle
import random

import numpy as np
import pytest

from truepix_amd import native_fp2_bridge as nb
from truepix_amd.extfield import _raw, proof_pickle_names
from truepix_amd.fiatshamir import FiatShamir, native_dumps

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = (1 << 61) - 1


@pytest.fixture(scope="module")
def lib():
    return nb.load_library()


def _sha(lib, data, chunk):
    out = ctypes.create_string_buffer(32)
    lib.tp_sha256(data, len(data), chunk, out)
    return out.raw


def test_host_sha256_matches_hashlib(lib):
    rng = random.Random(3)
    for n in [0, 1, 55, 56, 63, 64, 65, 119, 120, 127, 128, 1000, 4096, 100003]:
        data = bytes(rng.getrandbits(8) for _ in range(n))
        for chunk in (0, 1, 7, 64, 100):
            assert _sha(lib, data, chunk) == hashlib.sha256(data).digest(), (n, chunk)
    big = np.random.default_rng(1).integers(0, 2 ** 63, size=1 << 18, dtype=np.uint64).tobytes()
    assert _sha(lib, big, 0) == hashlib.sha256(big).digest()
    assert _sha(lib, b"abc", 0).hex() == \
        "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"   # FIPS 180-4 KAT


def test_proof_writer_reproduces_golden_proofs(lib):
    with open(os.path.join(GOLD, "golden_proofs.json")) as fh:
        idx = json.load(fh)
    for e in idx["cases"]:
        if "proof_file" not in e:
            continue
        with open(os.path.join(GOLD, e["proof_file"]), "rb") as fh:
            raw = bz2.decompress(fh.read())
        (q, io, trans, ndb, rvs, rve, wdiv) = FiatShamir.unpack_proof(raw)
        assert native_dumps((q, list(io), list(trans), ndb, rvs, rve, wdiv, 4), lib) == raw, e["name"]


def _rint(r):
    return r.getrandbits(r.choice([8, 16, 31, 32, 40, 56, 61, 63, 64]))


def _leaf(r):
    k = r.random()
    if k < 0.4:
        return _rint(r)
    if k < 0.6:
        return _raw(r.randrange(P), r.randrange(P)) if r.random() < 0.8 else \
            _raw(r.randrange(300), r.randrange(70000))
    if k < 0.8:
        return "s%d" % r.getrandbits(20) + ("x" * r.choice([0, 10, 300]))
    return None


def _tree(r, d):
    c = r.random()
    if d >= 2 or c < 0.3:
        return _leaf(r)
    if c < 0.5:
        return tuple(_tree(r, d + 1) for _ in range(r.choice([0, 1, 2, 3, 4, 7])))
    if c < 0.65:
        return [_rint(r) for _ in range(r.choice([2, 999, 1000, 1001, 2001, 40000, 70001]))]
    n = r.choice([0, 1, 2, 5, 1000, 1001]) if d == 0 else r.choice([0, 1, 2, 5])
    return [_tree(r, d + 1) for _ in range(n)]


def test_proof_writer_matches_pickle_on_random_trees(lib):
    """Frames of 64 KiB, batches of 1000, one-element lists, LONG1 widths,
    memo indices past 255: bytes equal pickle.dumps(obj, 5)."""
    for seed in range(25):
        r = random.Random(seed)
        obj = tuple(_tree(r, 0) for _ in range(8))
        with proof_pickle_names():
            want = pickle.dumps(obj, -1)
        assert native_dumps(obj, lib) == want, seed


def test_proof_writer_rejects_malformed_tokens(lib):
    ptr = ctypes.POINTER(ctypes.c_uint8)()
    n = ctypes.c_size_t()
    for bad in (b"", b"l\x02\x00\x00\x00i", b"q", b"s\xff\x00\x00\x00ab", b"nn"):
        assert lib.tp_pickle_tree(bad, len(bad), ctypes.byref(ptr), ctypes.byref(n)) == -1
