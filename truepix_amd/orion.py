"""The signer's Orion commitment of a frame on the device.

The reference's commit phase is a program, ``multivariate_PC_test_commit``
(OZ/examples/multivariate_PC_test_commit.cpp:265-331): it packs the frame's
input.txt into the copy-major-384-const128-pad512 layout, masks it with a
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


def commit(inputs, seed: int, timing: bool = False):
    """(masked_root, mask_root) hex strings of the signer's commitment of
    inputs [copies][384] (sample values 0..255) with the mask seed; with
    timing, also the device milliseconds."""
    nb.require()
    arr = np.ascontiguousarray(inputs)
    if arr.ndim != 2 or arr.shape[1] != COPY_SIZE:
        raise ValueError("inputs must be [copies][%d]" % COPY_SIZE)
    if arr.size and (arr.min() < 0 or arr.max() > 255):
        raise ValueError("sample values must be in 0..255")
    data = np.ascontiguousarray(arr, dtype=np.uint8)
    a = ctypes.create_string_buffer(65)
    b = ctypes.create_string_buffer(65)
    ms = ctypes.c_double()
    rc = nb.LIB.tp_orion_commit(int(arr.shape[0]), data.tobytes(), int(seed) & (2 ** 64 - 1), a, b,
                                ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError("tp_orion_commit failed: %s" % nb.last_error())
    roots = (a.value.decode(), b.value.decode())
    return roots + (ms.value,) if timing else roots


def commit_yuv(y, cb, cr, seed: int, timing: bool = False):
    """``commit`` of a frame given as YUV 4:2:0 planes, its per-copy rows
    built on the device (``yuv.macroblocks``)."""
    from truepix_amd import yuv
    return commit(yuv.macroblocks(y, cb, cr), seed, timing)


def manifest(inputs, seed: int, zk: bool = True) -> dict:
    """The orion_commit.json the reference's commit phase writes for these
    inputs and mask seed (truepix_gkr.h:428-455)."""
    copies = int(np.asarray(inputs).shape[0])
    masked_root, mask_root = commit(inputs, seed)
    padded = 1 << max(0, (copies - 1).bit_length())
    n = padded * 512
    return {"protocol": PROTOCOL_ID, "layout": LAYOUT_ID, "zk": zk, "copies": copies,
            "padded_copies": padded, "copy_size": COPY_SIZE, "padded_copy_size": 512,
            "const_idx": 384, "const_val": 128, "n": n, "lg_n": n.bit_length() - 1,
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
