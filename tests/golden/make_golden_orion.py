#!/usr/bin/env python3
"""Writes tests/golden/orion_ref.json from the reference's own commitment
pieces (the layout and mask of OZ/include/truepix/truepix_gkr.h, expander_init
and encode of OZ/include/linear_code/, the row hash and create_tree of
OZ/src/infrastructure/merkle_tree.cpp over flo-shani's SHA-256, compiled in
place by `make -C oracle ref` and driven in commit_pair's order by
oracle/ref_harness_orion.cpp): the masked and mask roots and the codeword
length for seeded inputs, SHA-256 digests of single-column codewords, and
SHA-256 digests of opening proofs (OpeningProof bytes from orion_proof.cpp's
serializer) at seeded points and queries, and digests of the Virgo
code-switching transcript that closes each NIZK opening (generate_circuit,
zk_prover and zk_verifier's checks, oracle/ref_harness_virgo.cpp).
Run where /root/reference exists:

    make -C oracle ref && python tests/golden/make_golden_orion.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

# (copies, mask seed, input seed): 1 and 3 copies keep every column below the
# code's distance threshold; the rest cover 1 .. 5 expander levels, up to the
# 1350-copy SD frame of the signer (n = 2^20).
COMMIT_CASES = [(1, 5, 1), (3, 7, 2), (5, 11, 3), (16, 2024, 4), (33, 99, 5), (300, 12345, 6),
                (1350, 777, 7)]
ENCODE_SIZES = [5, 13, 14, 31, 64, 100, 1024, 8192]
# openings (tensor_product_protocol, with orion_proof.cpp's siblings and
# serializer): (copies, mask seed, input seed, use_group1, point seed,
# queries); the point and the queries from PCG64(point seed)
OPEN_CASES = [(1, 5, 1, 1, 21, 8), (3, 7, 2, 0, 22, 16), (16, 2024, 4, 1, 23, 64),
              (300, 12345, 6, 0, 24, 200), (1350, 777, 7, 1, 25, 1222)]
# the other circuits' packings (copy_size, const_idx, const_val), as
# TP/libTruePix/fp2_link.py:533-563 assigns them: commit roots, the NIZK
# opening's proof bytes and its code-switching transcript, per (name, packing,
# copies, mask seed, input seed, use_group1, point seed)
LAYOUT_CASES = [("inv", (384, 384, 255), 16, 31, 8, 1, 41),
                ("mask", (384, -1, 0), 40, 32, 9, 0, 42),
                ("crop", (385, -1, 0), 300, 33, 10, 1, 43)]


def inputs_for(copies, seed, copy_size=384):
    return np.random.Generator(np.random.PCG64(seed)).integers(0, 256, size=(copies, copy_size),
                                                               dtype=np.uint8)


def column_for(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.integers(0, (1 << 61) - 1, size=(n, 2), dtype=np.uint64)


def open_args(copies, point_seed, nq, cw_len):
    """The point r (lg n challenges, (re, im)) and nq code positions."""
    pc = 1
    while pc < copies:
        pc <<= 1
    lg = (pc * 512).bit_length() - 1
    rng = np.random.Generator(np.random.PCG64(point_seed))
    r = rng.integers(0, (1 << 61) - 1, size=(lg, 2), dtype=np.uint64)
    q = rng.integers(0, cw_len, size=nq, dtype=np.int64)
    q[:2] = (0, cw_len - 1)
    return r, q


def root_bytes(hex_root: str) -> bytes:
    """A digest's bytes from its root_hex form (its u64 words at byte offsets
    8, 0, 24, 16, each printed as 16 hex digits)."""
    words = [int(hex_root[16 * i:16 * i + 16], 16) for i in range(4)]
    out = bytearray(32)
    for w, off in zip(words, (8, 0, 24, 16)):
        out[off:off + 8] = w.to_bytes(8, "little")
    return bytes(out)


def opening_fields(proof: bytes) -> dict:
    """The opened and combined roots and the combined message of an
    OpeningProof version-2 byte string (OZ/include/truepix/orion_proof.h)."""
    import struct
    n, cols, msg_len, cw_len = struct.unpack_from("<QQQQ", proof, 16)
    (nq,) = struct.unpack_from("<I", proof, 48)
    o = 52 + 16
    com_root, comb_root = proof[o:o + 32], proof[o + 32:o + 64]
    o += 64 + 8 * nq + 16 * nq * cols + 16 * cw_len
    fe = np.frombuffer(proof, dtype="<u8", count=2 * msg_len, offset=o).reshape(msg_len, 2)
    return {"com_root": com_root, "combined_root": comb_root,
            "combined_message": fe[:, ::-1].copy()}   # (imag, real) on disk


def main():
    from oracle.oracle import RefOrion
    if not RefOrion.available():
        raise SystemExit("build the reference harness first: make -C oracle ref")
    ref = RefOrion()
    out = {"source": "truepix_gkr.h layout + mask, expander_init/encode, merkle_tree.cpp "
                     "(commit_pair order), glibc rand() initial state",
           "inputs": "PCG64(input_seed).integers(0, 256, (copies, 384), uint8)",
           "columns": "PCG64(n).integers(0, 2^61 - 1, (n, 2), uint64) as (re, im)",
           "commit": [], "encode": []}
    for copies, seed, iseed in COMMIT_CASES:
        m, k, cw = ref.commit(inputs_for(copies, iseed), seed)
        out["commit"].append({"copies": copies, "seed": seed, "input_seed": iseed,
                              "masked_root": m, "mask_root": k, "cw_len": cw})
    for n in ENCODE_SIZES:
        cwd = ref.encode(column_for(n, n))
        out["encode"].append({"n": n, "cw_len": int(cwd.shape[0]),
                              "sha256": hashlib.sha256(cwd.tobytes()).hexdigest(),
                              "head": [int(x) for x in cwd[:4].reshape(-1)]})
    out["open"] = []
    cw_of = {c["copies"]: c["cw_len"] for c in out["commit"]}
    for copies, seed, iseed, group, pseed, nq in OPEN_CASES:
        r, q = open_args(copies, pseed, nq, cw_of[copies])
        proof = ref.open(inputs_for(copies, iseed), seed, r, q, use_group1=bool(group))
        out["open"].append({"copies": copies, "seed": seed, "input_seed": iseed,
                            "use_group1": group, "point_seed": pseed, "queries": nq,
                            "proof_len": len(proof),
                            "proof_sha256": hashlib.sha256(proof).hexdigest()})
    # the NIZK query draw (orion_fs.cpp's transcript) for the openings' roots
    out["draw"] = []
    for copies, seed, iseed, group, pseed, nq in OPEN_CASES:
        r, _ = open_args(copies, pseed, nq, cw_of[copies])
        m, k, cw = ref.commit(inputs_for(copies, iseed), seed)
        root = root_bytes(m if group else k)
        pc = 1
        while pc < copies:
            pc <<= 1
        q = ref.draw_queries(root, pc * 512, bool(group), cw, r)
        out["draw"].append({"copies": copies, "seed": seed, "input_seed": iseed,
                            "use_group1": group, "point_seed": pseed, "queries": len(q),
                            "head": q[:8], "sha256": hashlib.sha256(np.asarray(q, "<i8").tobytes()).hexdigest()})
    # the Virgo code-switching proof of each NIZK opening (queries drawn as
    # above): the transcript words of oracle/ref_harness_virgo.cpp
    from oracle.oracle import RefVirgo
    virgo = RefVirgo()
    out["virgo"] = []
    for copies, seed, iseed, group, pseed, nq in OPEN_CASES:
        r, _ = open_args(copies, pseed, nq, cw_of[copies])
        m, k, cw = ref.commit(inputs_for(copies, iseed), seed)
        pc = 1
        while pc < copies:
            pc <<= 1
        q = ref.draw_queries(root_bytes(m if group else k), pc * 512, bool(group), cw, r)
        proof = ref.open(inputs_for(copies, iseed), seed, r, np.asarray(q, np.int64),
                         use_group1=bool(group))
        f = opening_fields(proof)
        words = virgo.transcript(f["com_root"], pc * 512, bool(group), r, f["combined_root"],
                                 f["combined_message"])
        depth = int(words[0])
        out["virgo"].append({"copies": copies, "seed": seed, "input_seed": iseed,
                             "use_group1": group, "point_seed": pseed, "words": int(words.size),
                             "bits": [int(x) for x in words[1:1 + depth]],
                             "verdict": int(words[-1]),
                             "tail": [int(x) for x in words[-5:]],
                             "sha256": hashlib.sha256(words.astype("<u8").tobytes()).hexdigest()})
    # the other packings, through the environment the reference reads
    from oracle.oracle import RefOrion as _R
    out["layouts"] = []
    for name, lay, copies, seed, iseed, group, pseed in LAYOUT_CASES:
        _R.set_layout(*lay)
        try:
            x = inputs_for(copies, iseed, lay[0])
            m, k, cw = ref.commit(x, seed)
            r, _ = open_args(copies, pseed, 2, cw)
            pc = 1 << max(0, (copies - 1).bit_length())
            q = ref.draw_queries(root_bytes(m if group else k), pc * 512, bool(group), cw, r)
            proof = ref.open(x, seed, r, np.asarray(q, np.int64), use_group1=bool(group))
            f = opening_fields(proof)
            words = virgo.transcript(f["com_root"], pc * 512, bool(group), r, f["combined_root"],
                                     f["combined_message"])
        finally:
            _R.set_layout(384, 384, 128)
        out["layouts"].append({"name": name, "layout": list(lay), "copies": copies, "seed": seed,
                               "input_seed": iseed, "use_group1": group, "point_seed": pseed,
                               "masked_root": m, "mask_root": k, "cw_len": cw,
                               "proof_len": len(proof),
                               "proof_sha256": hashlib.sha256(proof).hexdigest(),
                               "virgo_words": int(words.size), "virgo_verdict": int(words[-1]),
                               "virgo_sha256": hashlib.sha256(words.astype("<u8").tobytes()).hexdigest()})
    with open(os.path.join(HERE, "orion_ref.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", len(out["commit"]), "commit cases,", len(out["encode"]), "encode cases,",
          len(out["open"]), "openings")


if __name__ == "__main__":
    main()
