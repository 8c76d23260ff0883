"""Commitment-manifest binding, point files and random circuit I/O for the
truepix protocol (subset of TP/libTruePix/fp2_link.py used by the GKR path).

  commit_binding(manifest)  the list absorbed into the transcript before the
                            GKR challenges (TP/libTruePix/fp2_link.py:487-504)
  read_commit_manifest()    orion_commit.json, validated as the reference
                            validates it (:426-484)
  write_point_file()        GKR input point for the Orion opening (:63-79)
  prepare_random_circuit_io()  random uint8 macroblocks -> input.txt /
                            output.txt (:302-400), here from a seeded PCG64
                            generator (or the OS CSPRNG) and evaluated on the GPU
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from truepix_amd.extfield import BASE_PRIME_61, Fp2, deserialize_point, serialize_point

LAYOUT_ID = "copy-major-384-const128-pad512-v1"
PROTOCOL_ID = "truepix"
ORION_COMMIT_FILE = "orion_commit.json"
_PROTOCOL_ALIASES = {PROTOCOL_ID, "fp2-linked-v1"}
_MANIFEST_INT_FIELDS = ("copies", "padded_copies", "copy_size", "padded_copy_size",
                        "const_idx", "const_val", "n", "lg_n")


@dataclass
class PendingInputClaim:
    input_point: List[Fp2]
    authenticated_claim: object
    n_copies: int
    n_in_bits: int


def write_point_file(path: str, point: Sequence[Fp2], layout: str = LAYOUT_ID):
    with open(path, "w") as fh:
        json.dump({"protocol": PROTOCOL_ID, "field": "fp2-mersenne61", "layout": layout,
                   "point": serialize_point(point)}, fh, indent=2)
    txt = path if path.endswith(".txt") else os.path.splitext(path)[0] + ".txt"
    with open(txt, "w") as fh:
        fh.write("%d\n" % len(point))
        for e in point:
            fh.write("%d %d\n" % (e.real, e.imag))
    return path, txt


def read_point_file(path: str) -> List[Fp2]:
    with open(path) as fh:
        data = json.load(fh)
    if data.get("protocol") not in _PROTOCOL_ALIASES:
        raise ValueError("point file protocol mismatch")
    return deserialize_point(data["point"], BASE_PRIME_61)


def _check_root(manifest, key, path, required):
    v = manifest.get(key)
    if required or v:
        if not isinstance(v, str) or len(v) != 64:
            raise ValueError("commitment manifest %s has a bad %s" % (path, key))
        int(v, 16)


def read_commit_manifest(path: Optional[str] = None) -> dict:
    if path is None:
        path = os.environ.get("TRUEPIX_ORION_COMMIT_FILE", ORION_COMMIT_FILE)
    if not os.path.isfile(path):
        raise IOError("missing Orion commitment manifest %s; run the commit phase before "
                      "proving" % path)
    with open(path) as fh:
        manifest = json.load(fh)
    if manifest.get("protocol") not in _PROTOCOL_ALIASES:
        raise ValueError("commitment manifest %s has protocol %r" % (path, manifest.get("protocol")))
    if manifest.get("layout") != LAYOUT_ID:
        raise ValueError("commitment manifest %s has layout %r, expected %r"
                         % (path, manifest.get("layout"), LAYOUT_ID))
    for field in _MANIFEST_INT_FIELDS:
        if field not in manifest:
            raise ValueError("commitment manifest %s is missing %r" % (path, field))
        manifest[field] = int(manifest[field])
    zk = manifest.get("zk", True)
    if isinstance(zk, str):
        zk = zk.strip().lower() in ("1", "true", "yes")
    manifest["zk"] = bool(zk)
    _check_root(manifest, "masked_root", path, True)
    mask = manifest.get("mask_root", "")
    if mask is None:
        mask = ""
    if not isinstance(mask, str):
        raise ValueError("commitment manifest %s has a bad mask_root" % path)
    manifest["mask_root"] = mask
    _check_root(manifest, "mask_root", path, manifest["zk"])
    return manifest


def commit_binding(manifest: dict) -> list:
    return (["orion-commit", PROTOCOL_ID, LAYOUT_ID, 1 if manifest.get("zk", True) else 0]
            + [int(manifest[f]) for f in _MANIFEST_INT_FIELDS]
            + [manifest["masked_root"], manifest.get("mask_root") or ""])


def make_manifest(n_copies: int, copy_size: int = 384, const_idx: int = 384, const_val: int = 128,
                  seed: int = 0, zk: bool = True) -> dict:
    """A well-formed manifest with deterministic pseudo-random roots (for tests and
    benchmarks that run without the Orion commit binaries)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    padded = 1 << max(0, (n_copies - 1).bit_length())
    n = padded * 512
    roots = [rng.bytes(32).hex() for _ in range(2)]
    return {"protocol": PROTOCOL_ID, "layout": LAYOUT_ID, "zk": zk, "copies": n_copies,
            "padded_copies": padded, "copy_size": copy_size, "padded_copy_size": 512,
            "const_idx": const_idx, "const_val": const_val, "n": n, "lg_n": n.bit_length() - 1,
            "masked_root": roots[0], "mask_root": roots[1] if zk else ""}


# Input packing per PWS file name prefix: (copy_size, const_idx, const_val), as
# TP/libTruePix/fp2_link.py:533-563 (infer_gkr_input_layout) assigns them; a
# name with no entry keeps the gray packing.
_PWS_LAYOUTS = (("inv", (384, 384, 255)), ("mask", (384, -1, 0)), ("crop", (385, -1, 0)))
_DEFAULT_LAYOUT = (384, 384, 128)


def infer_gkr_input_layout(pws_path: Optional[str] = None, input_path: str = "input.txt"):
    """(copy_size, const_idx, const_val) of the private-input packing the Orion
    commitment is made over, from the PWS file's name; an existing input.txt
    fixes copy_size to the length of its first line (fp2_link.py:533-563)."""
    layout = _DEFAULT_LAYOUT
    if pws_path:
        base = os.path.basename(pws_path).lower()
        for prefix, lay in _PWS_LAYOUTS:
            if base.startswith(prefix):
                layout = lay
                break
    copy_size, const_idx, const_val = layout
    if os.path.isfile(input_path):
        with open(input_path) as fh:
            first = fh.readline().split()
        if first:
            copy_size = len(first)
    return copy_size, const_idx, const_val


def input_layout(parsed: list):
    """(copy_size, const_idx, const_val) read off a parsed circuit's input
    layer: the private inputs per copy, and the wire index and value of its
    one constant wire (-1, 0 without one)."""
    layer = parsed[0]
    consts = [(i, int(c)) for i, c in enumerate(layer) if c is not None]
    if len(consts) > 1:
        raise ValueError("the truepix packing has at most one constant wire (got %d)" % len(consts))
    idx, val = consts[0] if consts else (-1, 0)
    return sum(1 for c in layer if c is None), idx, val


def write_manifest(path: str, manifest: dict) -> None:
    with open(path, "w") as fh:
        json.dump(manifest, fh, indent=2)


def random_private_inputs(circuit_name: str, parsed: list, n_copies: int, seed: Optional[int] = None,
                          value_max: int = 255, crop_region=(0.2, 0.8, 0.2, 0.8),
                          mb_width: Optional[int] = None) -> np.ndarray:
    """[n_copies, n_private] uint8-range private inputs.  For crop the last
    private wire is the keep bit over a rectangle of macroblocks
    (TP/libTruePix/fp2_link.py:270-300)."""
    n_priv = sum(1 for x in parsed[0] if x is None)
    if seed is None:
        rng = np.random.Generator(np.random.PCG64(int.from_bytes(os.urandom(16), "little")))
    else:
        rng = np.random.Generator(np.random.PCG64(seed))
    vals = rng.integers(0, value_max + 1, size=(n_copies, n_priv), dtype=np.int64)
    if circuit_name.lower().startswith("crop"):
        vals[:, -1] = crop_bits(n_copies, *crop_region, mb_width=mb_width)
    return vals


def crop_bits(n_copies, x0, x1, y0, y1, mb_width=None):
    if mb_width is None:
        mb_width = int(round((n_copies * 4 / 3) ** 0.5))     # 4:3 frame by default
        mb_width = max(1, mb_width)
    h = (n_copies + mb_width - 1) // mb_width
    bits = np.zeros(n_copies, dtype=np.int64)
    xs, xe = int(x0 * mb_width), int(x1 * mb_width)
    ys, ye = int(y0 * h), int(y1 * h)
    for y in range(ys, ye):
        for x in range(xs, xe):
            if y * mb_width + x < n_copies:
                bits[y * mb_width + x] = 1
    return bits


def full_input_rows(parsed: list, private: np.ndarray) -> np.ndarray:
    """Insert constant wires: [n_copies, len(input_layer)] values (util.get_inputs)."""
    layer = parsed[0]
    rows = np.zeros((private.shape[0], len(layer)), dtype=np.int64)
    k = 0
    for i, c in enumerate(layer):
        if c is None:
            rows[:, i] = private[:, k]
            k += 1
        else:
            rows[:, i] = int(c) % BASE_PRIME_61
    return rows


def write_io_files(private: np.ndarray, outputs: np.ndarray, input_path: str = "input.txt",
                   output_path: str = "output.txt", n_public: int = 384) -> None:
    with open(input_path, "w") as fh:
        for row in private.tolist():
            fh.write(" ".join(str(v) for v in row) + "\n")
    with open(output_path, "w") as fh:
        for row in outputs[:, :n_public].tolist():
            fh.write(" ".join(str(int(v) % BASE_PRIME_61) for v in row) + "\n")
