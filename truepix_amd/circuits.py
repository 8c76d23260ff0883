"""Layered-circuit descriptions for the TruePix edit circuits.

A parsed circuit has the shape pypws returns and TP/libTruePix/parse_pws.py:8-64
consumes: ``[input_layer, layer_1, ..., layer_L]`` where ``input_layer`` lists
``None`` for each private input and an int for each constant wire, and each
layer is a list of ``(op, in0, in1, mux)`` tuples (op in ADD, MUL, SUB, MUX,
OR, XOR, NOT, NAND, NOR, NXOR, NAAB, PASS).

The four circuits of TP/pws/ are restated here from their meaning (one
macroblock = 256 Y + 64 Cb + 64 Cr samples per copy):
  gray    Y kept, Cb = Cr = 128          (TP/pws/gray.pws; VideoEdition/Gray.py:7-10)
  invert  255 - x for every sample       (TP/pws/inv.pws;  VideoEdition/Inv.py:4-9)
  crop    x * keep_bit, keep_bit = input 384 (TP/pws/Crop.pws; its constant 128 is
          unused and the PWS optimizer drops it, TP/libTruePix/fp2_link.py:547-552)
  mask    every Y sample = Y[136], Cb = Cb[36] (wire 292), Cr = Cr[36] (wire 356)
          (TP/pws/mask.pws)
``parse_pws_text`` reads the PWS subset those files use, with the optimizer
steps that matter for them (unused-input removal and duplicate-gate removal,
R/libpws/src/circuit/pws_circuit_parser.cc:1527-1773).
"""
from __future__ import annotations

import re
from typing import List, Optional, Tuple

Gate = Tuple[str, int, int, Optional[int]]

OPS = ("ADD", "MUL", "SUB", "MUX", "OR", "XOR", "NOT", "NAND", "NOR", "NXOR", "NAAB", "PASS")
# type index per op (gate_type_idx of the reference gate classes)
OP_INDEX = {"MUL": 0, "ADD": 1, "SUB": 2, "MUX": 3, "OR": 5, "XOR": 6, "NOT": 7,
            "NAND": 8, "NOR": 9, "NXOR": 10, "NAAB": 11, "PASS": 12}

N_SAMPLES = 384          # 256 Y + 64 Cb + 64 Cr per 16x16 4:2:0 macroblock


def gray() -> list:
    inputs = [None] * N_SAMPLES + [128]
    gates = [("PASS", i, i, None) for i in range(256)]
    gates += [("PASS", N_SAMPLES, 256 + j, None) for j in range(128)]
    return [inputs, gates]


def invert() -> list:
    inputs = [None] * N_SAMPLES + [255]
    return [inputs, [("SUB", N_SAMPLES, i, None) for i in range(N_SAMPLES)]]


def crop() -> list:
    inputs = [None] * (N_SAMPLES + 1)           # samples + the keep bit
    return [inputs, [("MUL", i, N_SAMPLES, None) for i in range(N_SAMPLES)]]


def mask() -> list:
    inputs = [None] * N_SAMPLES
    gates = [("PASS", 136, i, None) for i in range(256)]
    gates += [("PASS", 292, i, None) for i in range(256, 320)]
    gates += [("PASS", 356, i, None) for i in range(320, 384)]
    return [inputs, gates]


CIRCUITS = {"gray": gray, "invert": invert, "inv": invert, "crop": crop, "mask": mask}


def circuit(name: str) -> list:
    try:
        return CIRCUITS[name.lower()]()
    except KeyError:
        raise ValueError("unknown circuit %r (have %s)" % (name, sorted(CIRCUITS)))


# --- PWS subset ---------------------------------------------------------------

_PWS_OPS = {"PASS": "PASS", "minus": "SUB", "-": "SUB", "*": "MUL", "+": "ADD"}
_LINE = re.compile(r"^P\s+([VO])(\d+)\s*=\s*(.*?)\s+E\s*$")


def parse_pws_text(text: str, optimize: bool = True) -> list:
    """Parse a single-layer PWS program of the form used by TP/pws/*.pws:
    ``P Vk = Ij E`` (private input), ``P Vk = <int> E`` (constant) and
    ``P Ok = Va <op> Vb E`` with op in PASS, minus, *, +."""
    wires = {}            # V name -> ("in", j) | ("const", value)
    order: List[int] = []
    outs: List[Tuple[int, str, int, int]] = []
    for raw in text.splitlines():
        line = raw.strip()
        if not line:
            continue
        m = _LINE.match(line)
        if not m:
            raise ValueError("unsupported PWS line: %r" % raw)
        kind, num, rhs = m.group(1), int(m.group(2)), m.group(3).split()
        if kind == "V":
            if len(rhs) != 1:
                raise ValueError("unsupported PWS assignment: %r" % raw)
            if rhs[0].startswith("I"):
                wires[num] = ("in", int(rhs[0][1:]))
            else:
                wires[num] = ("const", int(rhs[0]))
            order.append(num)
        else:
            if len(rhs) != 3 or rhs[1] not in _PWS_OPS:
                raise ValueError("unsupported PWS gate: %r" % raw)
            a, b = rhs[0], rhs[2]
            if not (a.startswith("V") and b.startswith("V")):
                raise ValueError("unsupported PWS operand: %r" % raw)
            outs.append((num, _PWS_OPS[rhs[1]], int(a[1:]), int(b[1:])))
    # input layer: wires in order of definition
    pos = {v: i for i, v in enumerate(order)}
    input_layer = [None if wires[v][0] == "in" else wires[v][1] for v in order]
    gates = [(op, pos[a], pos[b], None) for (_, op, a, b) in sorted(outs, key=lambda t: t[0])]
    if optimize:
        used = sorted({g[1] for g in gates} | {g[2] for g in gates})
        remap = {old: new for new, old in enumerate(used)}
        input_layer = [input_layer[i] for i in used]
        gates = [(op, remap[a], remap[b], mx) for (op, a, b, mx) in gates]
        seen = set()
        dedup = []
        for g in gates:                 # output layer: duplicates are removed
            if g in seen:
                continue
            seen.add(g)
            dedup.append(g)
        gates = dedup
    return [input_layer, gates]


def parse_pws_file(path: str, optimize: bool = True) -> list:
    with open(path) as fh:
        return parse_pws_text(fh.read(), optimize)


def n_private(parsed: list) -> int:
    return sum(1 for x in parsed[0] if x is None)
