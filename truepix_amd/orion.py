"""The signer's Orion commitment of a frame on the device.

The reference's commit phase is a program, ``multivariate_PC_test_commit``
(OZ/examples/multivariate_PC_test_commit.cpp:265-331): it packs the frame's
input.txt into the copy-major pad-512 layout with the circuit's packing
(``layout``: copy_size samples and a constant wire; gray 384/384/128,
invert 384/384/255, mask 384/-1/0, crop 385/-1/0), masks it with a
seeded mask (OZ/include/truepix/truepix_gkr.h:286-386), commits the masked and
the mask polynomials with commit_pair (OZ/src/VPD/linearPC.cpp:253-459: the
expander code per column, SHA-256 row chains, a Merkle tree over the rows) and
writes the two roots into orion_commit.json (truepix_gkr.h:428-455).  The
frame prover binds its transcript to that manifest
(``fp2_link.commit_binding``).

``commit(inputs, seed)`` computes the two roots in ``libtruepix_hip.so``
(``tp_orion_commit``); ``manifest(inputs, seed)`` wraps them in the manifest
the reference writes; ``encode(columns)`` exposes the expander code alone.
"""
from __future__ import annotations

import ctypes
from typing import Tuple

import numpy as np

from truepix_amd import native_fp2_bridge as nb
from truepix_amd.fp2_link import LAYOUT_ID, PROTOCOL_ID

COPY_SIZE = 384
_U64P = ctypes.POINTER(ctypes.c_uint64)


GRAY_LAYOUT = (384, 384, 128)   # (copy_size, const_idx, const_val): the gray circuit's packing


def _samples(inputs, layout):
    """The frame's samples as bytes, with the library set to the packing
    (copy_size samples per copy, the constant const_val at wire const_idx or
    none when -1; truepix_gkr.h:107-150).  The packing is process-wide in the
    library, as the reference's programs take it from their environment."""
    copy_size, const_idx, const_val = (int(x) for x in layout)
    arr = np.ascontiguousarray(inputs)
    if arr.ndim != 2 or arr.shape[1] != copy_size:
        raise ValueError("inputs must be [copies][%d]" % copy_size)
    if arr.size and (arr.min() < 0 or arr.max() > 255):
        raise ValueError("sample values must be in 0..255")
    if nb.LIB.tp_orion_set_layout(copy_size, const_idx, const_val) != 0:
        raise ValueError(nb.last_error())
    return int(arr.shape[0]), np.ascontiguousarray(arr, dtype=np.uint8).tobytes()


def commit(inputs, seed: int, timing: bool = False, layout=GRAY_LAYOUT):
    """(masked_root, mask_root) hex strings of the signer's commitment of
    inputs [copies][copy_size] (sample values 0..255) with the mask seed and
    the circuit's packing; with timing, also the device milliseconds."""
    nb.require()
    copies, data = _samples(inputs, layout)
    a = ctypes.create_string_buffer(65)
    b = ctypes.create_string_buffer(65)
    ms = ctypes.c_double()
    rc = nb.LIB.tp_orion_commit(copies, data, int(seed) & (2 ** 64 - 1), a, b, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError("tp_orion_commit failed: %s" % nb.last_error())
    roots = (a.value.decode(), b.value.decode())
    return roots + (ms.value,) if timing else roots


def opening(inputs, seed: int, r, col_q, use_group1: bool = True, layout=GRAY_LAYOUT):
    """The opening of the committed masked polynomial (use_group1) or the mask
    at the point r (lg n challenges, each (re, im)), at the code positions
    col_q the transcript drew: (proof bytes in the reference's OpeningProof
    version-2 layout, the claimed value (re, im)).  OZ/src/VPD/linearPC.cpp:
    700-966 (tensor_product_protocol) up to its Virgo code-switching proof,
    which is not part of it."""
    nb.require()
    copies, data = _samples(inputs, layout)
    rr = np.ascontiguousarray(np.asarray(r, dtype=np.uint64).reshape(-1))
    q = np.ascontiguousarray(col_q, dtype=np.int64)
    buf = ctypes.c_void_p()
    n = ctypes.c_size_t()
    ans = np.zeros(2, np.uint64)
    rc = nb.LIB.tp_orion_open(copies, data, int(seed) & (2 ** 64 - 1),
                              1 if use_group1 else 0, rr.ctypes.data_as(_U64P), rr.size // 2,
                              q.ctypes.data, q.size, ctypes.byref(buf), ctypes.byref(n),
                              ans.ctypes.data_as(_U64P))
    if rc != 0:
        raise RuntimeError("tp_orion_open failed: %s" % nb.last_error())
    proof = ctypes.string_at(buf.value, n.value)
    nb.LIB.tp_free(buf)
    return proof, (int(ans[0]), int(ans[1]))


def opening_nizk(inputs, seed: int, r, use_group1: bool = True, layout=GRAY_LAYOUT):
    """``opening`` with the code positions drawn as the reference's NIZK mode
    draws them (OZ/src/truepix/orion_fs.cpp's transcript over the label
    "tensor-open", the opened root, N, the group flag, r0 and r1; 1222
    positions from the opened group, linearPC.cpp:711-758)."""
    nb.require()
    copies, data = _samples(inputs, layout)
    rr = np.ascontiguousarray(np.asarray(r, dtype=np.uint64).reshape(-1))
    buf = ctypes.c_void_p()
    n = ctypes.c_size_t()
    ans = np.zeros(2, np.uint64)
    rc = nb.LIB.tp_orion_open_nizk(copies, data, int(seed) & (2 ** 64 - 1),
                                   1 if use_group1 else 0, rr.ctypes.data_as(_U64P), rr.size // 2,
                                   ctypes.byref(buf), ctypes.byref(n), ans.ctypes.data_as(_U64P))
    if rc != 0:
        raise RuntimeError("tp_orion_open_nizk failed: %s" % nb.last_error())
    proof = ctypes.string_at(buf.value, n.value)
    nb.LIB.tp_free(buf)
    return proof, (int(ans[0]), int(ans[1]))


def opening_nizk_virgo(inputs, seed: int, r, use_group1: bool = True, layout=GRAY_LAYOUT):
    """``opening_nizk`` and the words of its Virgo code-switching transcript
    (layout in include/truepix_hip.h, tp_orion_open_nizk_virgo).  Returns
    (proof or None, answer, words): proof is None when the code-switching
    proof does not check, as the reference then writes no proof."""
    nb.require()
    copies, data = _samples(inputs, layout)
    rr = np.ascontiguousarray(np.asarray(r, dtype=np.uint64).reshape(-1))
    cap = 1 << 15
    while True:
        buf = ctypes.c_void_p()
        n = ctypes.c_size_t()
        ans = np.zeros(2, np.uint64)
        words = np.zeros(cap, np.uint64)
        nw = ctypes.c_longlong()
        rc = nb.LIB.tp_orion_open_nizk_virgo(
            copies, data, int(seed) & (2 ** 64 - 1), 1 if use_group1 else 0,
            rr.ctypes.data_as(_U64P), rr.size // 2, ctypes.byref(buf), ctypes.byref(n),
            ans.ctypes.data_as(_U64P), words.ctypes.data_as(_U64P), cap, ctypes.byref(nw))
        if buf.value:
            proof = ctypes.string_at(buf.value, n.value)
            nb.LIB.tp_free(buf)
        else:
            proof = None
        if nw.value > cap:
            cap = nw.value
            continue
        if rc != 0 and (nw.value == 0 or words[nw.value - 1] != 0):
            raise RuntimeError("tp_orion_open_nizk_virgo failed: %s" % nb.last_error())
        return (proof if rc == 0 else None), (int(ans[0]), int(ans[1])), words[:nw.value].copy()


def virgo_stats():
    """(device ms, launches, sumcheck layers) of the last code-switching proof."""
    ms, la, ly = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
    nb.LIB.tp_orion_virgo_stats(ctypes.byref(ms), ctypes.byref(la), ctypes.byref(ly))
    return ms.value, la.value, ly.value


def parse_opening(proof: bytes) -> dict:
    """The fields of an OpeningProof version-2 byte string
    (OZ/include/truepix/orion_proof.h:1-22); field elements as (re, im),
    digests as bytes."""
    import struct
    o = 0

    def take(fmt):
        nonlocal o
        v = struct.unpack_from("<" + fmt, proof, o)
        o += struct.calcsize("<" + fmt)
        return v

    def fes(k):
        nonlocal o
        a = np.frombuffer(proof, dtype="<u8", count=2 * k, offset=o).reshape(k, 2)
        o += 16 * k
        return a[:, ::-1].copy()   # (imag, real) on disk

    if proof[:8] != b"ORIONPV1":
        raise ValueError("not an OpeningProof")
    o = 8
    version, group = take("II")
    n, column_size, msg_len, cw_len = take("QQQQ")
    (nq,) = take("I")
    answer = tuple(int(x) for x in fes(1)[0])
    com_root, comb_root = proof[o:o + 32], proof[o + 32:o + 64]
    o += 64
    col_q = list(take("%dq" % nq)) if nq else []
    queried = fes(nq * column_size).reshape(nq, column_size, 2)
    combined_codeword = fes(cw_len)
    combined_message = fes(msg_len)

    def siblings():
        nonlocal o
        (k,) = take("Q")
        out = {}
        for _ in range(k):
            (i,) = take("I")
            out[i] = proof[o:o + 32]
            o += 32
        return out

    com_sib = siblings()
    comb_sib = siblings()
    if o != len(proof):
        raise ValueError("trailing bytes in the OpeningProof")
    return {"version": version, "use_group1": group, "N": n, "column_size": column_size,
            "msg_len": msg_len, "cw_len": cw_len, "answer": answer, "com_root": com_root,
            "combined_root": comb_root, "col_q": col_q, "queried_cols": queried,
            "combined_codeword": combined_codeword, "combined_message": combined_message,
            "com_siblings": com_sib, "combined_siblings": comb_sib}


def commit_yuv(y, cb, cr, seed: int, timing: bool = False, extra=None, layout=GRAY_LAYOUT):
    """``commit`` of a frame given as YUV 4:2:0 planes, its per-copy rows
    built on the device (``yuv.macroblocks``; ``extra`` appends the crop
    circuit's byte per copy, with its 385-sample packing)."""
    from truepix_amd import yuv
    return commit(yuv.macroblocks(y, cb, cr, extra), seed, timing, layout)


def manifest(inputs, seed: int, zk: bool = True, layout=GRAY_LAYOUT) -> dict:
    """The orion_commit.json the reference's commit phase writes for these
    inputs, mask seed and packing (truepix_gkr.h:428-455)."""
    copies = int(np.asarray(inputs).shape[0])
    masked_root, mask_root = commit(inputs, seed, layout=layout)
    padded = 1 << max(0, (copies - 1).bit_length())
    n = padded * 512
    copy_size, const_idx, const_val = (int(x) for x in layout)
    return {"protocol": PROTOCOL_ID, "layout": LAYOUT_ID, "zk": zk, "copies": copies,
            "padded_copies": padded, "copy_size": copy_size, "padded_copy_size": 512,
            "const_idx": const_idx, "const_val": const_val, "n": n, "lg_n": n.bit_length() - 1,
            "masked_root": masked_root, "mask_root": mask_root if zk else ""}


def encode(columns: np.ndarray) -> np.ndarray:
    """Codewords of columns [n_cols][n][2] (re, im u64), n_cols <= 256, with
    the expander graphs of length n: [n_cols][cw][2]."""
    nb.require()
    src = np.ascontiguousarray(columns, dtype=np.uint64)
    if src.ndim != 3 or src.shape[2] != 2:
        raise ValueError("columns must be [n_cols][n][2]")
    n_cols, n = src.shape[0], src.shape[1]
    # the codeword is at most 1.72 n long (n itself below 14)
    dst = np.zeros(n_cols * (int(n * 1.72) + 16) * 2, np.uint64)
    cw = nb.LIB.tp_orion_encode(n, n_cols, src.ctypes.data_as(_U64P), dst.ctypes.data_as(_U64P))
    if cw < 0:
        raise RuntimeError("tp_orion_encode failed: %s" % nb.last_error())
    return dst[:n_cols * cw * 2].reshape(n_cols, cw, 2).copy()
