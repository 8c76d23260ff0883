"""GKR prover and verifier for the truepix protocol, device-resident.

Python-facing API of the reference path (TP/libTruePix/circuitnizkvec_vole.py):
``ProverClass.from_pws(parsed_pws, nCopies) -> (input_layer, prover)``,
``prover.run(inputs) -> proof bytes``, ``VerifierClass.run(proof) ->
PendingInputClaim`` and ``verifier.finish_input(value)``.  The protocol steps
(transcript order, labels, VOLE commitments, equality / product proofs) follow
CircuitProverNIZK.run (:265-501) and CircuitVerifierNIZK.run (:758-958) so the
proof bytes are the reference's.  The arithmetic under them runs on the GPU
through libtruepix_hip.so:
  circuit evaluation (A3)          tp_eval_layer
  [wire][copy] tables (A4)         tp_wire_table
  early rounds (A6-A8)             early session (claim / fold / finish)
  beta tables (A5)                 native_fp2_compute_beta
  late rounds (A9)                 tp_late_round
  line projection values (A10)     tp_late_values
  wiring-predicate MLE (A11)       tp_wiring_mlext
  input / output MLE (A12)         tp_mle_eval_base_dev
Host code keeps the protocol glue the reference keeps on the host: round
polynomial interpolation, VOLE commitment arithmetic and the transcript.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence

import numpy as np

from truepix_amd import native_fp2_bridge as nb
from truepix_amd.extfield import P, Fp2, _as_fp2, _raw
from truepix_amd.fiatshamir import FiatShamir
from truepix_amd.fp2_link import PendingInputClaim, commit_binding, read_commit_manifest
from truepix_amd.extfield import serialize_point
from truepix_amd.vole import ExtVOLECommit

HALF = (P + 1) // 2           # inverse of 2 (TP/libTruePix/defs.py:18)
THIRD = (2 * P + 1) // 3      # inverse of 3 (defs.py:19)

OP_INDEX = {"MUL": 0, "ADD": 1, "SUB": 2, "MUX": 3, "OR": 5, "XOR": 6, "NOT": 7,
            "NAND": 8, "NOR": 9, "NXOR": 10, "NAAB": 11, "PASS": 12}


def clog2(v: int) -> int:
    return (int(v) - 1).bit_length()


# --- gate-function tables (reference: TP/libTruePix/gateprover.py:706-768 and
# gateprover_vole.py:690-752), indexed by gate type -------------------------
def _g(x, y, f):
    return f(x, y) % P


GATE_FN = [
    lambda x, y: (x * y) % P, lambda x, y: (x + y) % P, lambda x, y: (x - y) % P,
    lambda x, y: x, lambda x, y: y,
    lambda x, y: (x + y - x * y) % P, lambda x, y: (x + y - 2 * x * y) % P,
    lambda x, y: (1 - x) % P, lambda x, y: (1 - x * y) % P,
    lambda x, y: (1 + x * y - x - y) % P, lambda x, y: (1 + 2 * x * y - x - y) % P,
    lambda x, y: ((1 - x) * y) % P, lambda x, y: x,
]
# commitment-side forms f(x, y, xy): the prover combines tags linearly with the
# product tag standing in for xy (GateFunctionsPC), the verifier likewise
# (GateFunctionsVC).
GATE_COM_P = [
    lambda x, y, xy: xy, lambda x, y, xy: (x + y) % P, lambda x, y, xy: (x - y) % P,
    lambda x, y, xy: x, lambda x, y, xy: y,
    lambda x, y, xy: (x + y - xy) % P, lambda x, y, xy: (x + y - 2 * xy) % P,
    lambda x, y, xy: (1 - x) % P, lambda x, y, xy: (1 - xy) % P,
    lambda x, y, xy: (1 + xy - x - y) % P, lambda x, y, xy: (1 + 2 * xy - x - y) % P,
    lambda x, y, xy: (y - xy) % P, lambda x, y, xy: x,
]
GATE_COM_V = [
    lambda x, y, xy: xy, lambda x, y, xy: (x + y) % P, lambda x, y, xy: (x - y) % P,
    lambda x, y, xy: x, lambda x, y, xy: y,
    lambda x, y, xy: (x + y - x * y) % P, lambda x, y, xy: (x + y - 2 * x * y) % P,
    lambda x, y, xy: (1 - x) % P, lambda x, y, xy: (1 - xy) % P,
    lambda x, y, xy: (1 + xy - x - y) % P, lambda x, y, xy: (1 + 2 * xy - x - y) % P,
    lambda x, y, xy: (y - xy) % P, lambda x, y, xy: x,
]


# --- interpolation (TP/libTruePix/util.py:91-137, 279-312) ---------------------
def interpolate_quadratic(vals):
    """Coefficients [c0, c1, c2] of the quadratic through (0, f0), (1, f1), (-1, fm1)."""
    f0, f1, fm1 = vals
    c0 = f0
    c2 = ((f1 + fm1) * HALF - f0) % P
    c1 = (f1 - c2 - f0) % P
    return [c0 % P, c1, c2]


def interpolate_cubic(vals):
    """Coefficients [c0..c3] of the cubic through (0, f0), (1, f1), (-1, fm1), (2, f2)."""
    f0, f1, fm1, f2 = vals
    b = ((f1 + fm1) * HALF) % P          # c0 + c2
    c = ((f1 - fm1) * HALF) % P          # c1 + c3
    d = ((f2 - fm1) * THIRD) % P         # c1 + c2 + 3 c3
    c2 = (b - f0) % P
    c3 = ((d - c - c2) * HALF) % P
    c1 = (c - c3) % P
    return [f0 % P, c1, c2, c3]


def lagrange_coeffs(yvals):
    """Monomial coefficients of the polynomial through (t, yvals[t]), t = 0..n-1."""
    n = len(yvals)
    out = [0] * n
    for j in range(n):
        num = [1]
        den = 1
        for m in range(n):
            if m == j:
                continue
            nxt = [0] * (len(num) + 1)
            for i, a in enumerate(num):          # num * (x - m)
                nxt[i] = (nxt[i] - m * a) % P
                nxt[i + 1] = (nxt[i + 1] + a) % P
            num = nxt
            den = den * (j - m) % P
        scale = pow(den, P - 2, P)
        for i in range(n):
            coef = num[i] * scale % P
            if coef:
                out[i] = out[i] + yvals[j] * coef
    return [o % P if not isinstance(o, int) else o % P for o in out]


def horner_eval(coeffs, val):
    out = coeffs[-1]
    for e in reversed(coeffs[:-1]):
        out = (out * val + e) % P
    return out


# --- numpy / ctypes helpers ----------------------------------------------------
_U64P = ctypes.POINTER(ctypes.c_uint64)
_I32P = ctypes.POINTER(ctypes.c_int)


def _p(a: np.ndarray):
    return a.ctypes.data_as(_I32P if a.dtype == np.int32 else _U64P)


def _fp2_arr(vals: Sequence) -> np.ndarray:
    """Interleaved (re, im) uint64 array of Fp2 / int values."""
    out = np.zeros(2 * len(vals), dtype=np.uint64)
    for i, v in enumerate(vals):
        v = _as_fp2(v)
        out[2 * i], out[2 * i + 1] = v.real, v.imag
    return out


def _soa(vals: Sequence):
    a = _fp2_arr(vals)
    return np.ascontiguousarray(a[0::2]), np.ascontiguousarray(a[1::2])


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: %s" % (what, nb.last_error()))


def _torch():
    import torch
    return torch


# --- circuit description --------------------------------------------------------
class Layer:
    """One layer: gates reading the previous layer's 2^in_bits wires."""

    def __init__(self, in0, in1, ops, mux, in_bits):
        self.in0 = np.ascontiguousarray(in0, dtype=np.int32)
        self.in1 = np.ascontiguousarray(in1, dtype=np.int32)
        self.ops = list(ops)
        self.op_idx = np.ascontiguousarray([OP_INDEX[o] for o in ops], dtype=np.int32)
        self.mux = np.ascontiguousarray(mux, dtype=np.int32)
        self.n_gates = len(ops)
        self.out_bits = clog2(self.n_gates)
        self.in_bits = in_bits

    def type_idx(self, muxbits) -> np.ndarray:
        """Gate-type index with MUX resolved by its bit (native_fp2_bridge.py:450-472)."""
        t = self.op_idx.copy()
        for g in np.nonzero(t == 3)[0]:
            t[g] = 4 if muxbits[int(self.mux[g])] else 3
        return t


def parse_layers(parsed):
    """(input_layer, [Layer...]) from a pypws-shaped description
    (TP/libTruePix/parse_pws.py:8-64, including its padding of a one-gate
    output layer with an ADD gate)."""
    input_layer = parsed[0]
    layers = []
    in_bits = clog2(len(input_layer))
    specs = [list(l) for l in parsed[1:]]
    if len(specs[-1]) == 1:
        specs[-1].append(("ADD", 0, 0, None))
    for spec in specs:
        ops = [g[0] for g in spec]
        for o in ops:
            if o not in OP_INDEX:
                raise ValueError("Unknown gate type %s" % o)
        lay = Layer([g[1] for g in spec], [g[2] for g in spec], ops,
                    [g[3] if g[3] is not None else 0 for g in spec], in_bits)
        layers.append(lay)
        in_bits = lay.out_bits
    return input_layer, layers


# --- the prover engine ----------------------------------------------------------
class CircuitProver:
    """Device-resident layered GKR prover (TP/libTruePix/circuitprover.py,
    layerprover.py, layercompute.py, gateprover.py rewritten around device
    tables)."""

    def __init__(self, nCopies: int, nInputs: int, layers: List[Layer], muxbits=None):
        self.nCopies = nCopies
        self.nCopyBits = clog2(nCopies)
        self.nInputs = nInputs
        self.nInBits = clog2(nInputs)
        self.layers = layers
        self.muxlen = max([int(l.mux.max()) if l.n_gates else 0 for l in layers] + [0]) + 1
        self.muxbits = list(muxbits) if muxbits is not None else [0] * self.muxlen
        self.layerNum = 0
        self.values = []          # device rows per layer boundary: [copies_pad][2^bits]
        self.ckt_outputs = None   # host rows of the output layer

    # circuit evaluation (A3) --------------------------------------------------------
    def set_inputs(self, inputs) -> None:
        torch = _torch()
        assert len(inputs) == self.nCopies
        cp = 1 << self.nCopyBits
        rows = np.zeros((cp, 1 << self.nInBits), dtype=np.uint64)
        if isinstance(inputs, np.ndarray) and inputs.dtype.kind in "iu":
            arr = inputs.astype(np.uint64) % np.uint64(P)
        else:
            arr = np.array([[int(v) % P for v in row] for row in inputs], dtype=np.uint64)
        if arr.ndim != 2 or arr.shape[1] > rows.shape[1]:
            raise ValueError("inputs must be nCopies rows of at most 2^nInBits values")
        rows[:self.nCopies, :arr.shape[1]] = arr
        cur = torch.from_numpy(rows.view(np.int64)).to("cuda")
        self.values = [cur]
        stride = rows.shape[1]
        for lay in self.layers:
            out = torch.empty((cp, 1 << lay.out_bits), dtype=torch.int64, device="cuda")
            mux = np.ascontiguousarray([self.muxbits[m] if m < len(self.muxbits) else 0
                                        for m in lay.mux], dtype=np.int32)
            _check(nb.LIB.tp_eval_layer(ctypes.c_void_p(cur.data_ptr()), stride,
                                        ctypes.c_void_p(out.data_ptr()), lay.out_bits,
                                        self.nCopies, cp, _p(lay.op_idx), _p(lay.in0),
                                        _p(lay.in1), _p(mux), lay.n_gates, None), "tp_eval_layer")
            self.values.append(out)
            cur, stride = out, 1 << lay.out_bits
        self.ckt_outputs = cur.cpu().numpy().view(np.uint64)    # [copies_pad][2^out_bits]
        self.layerNum = 0
        self._begin_layer()

    def input_vector_device(self):
        return self.values[0]

    def _begin_layer(self):
        self.lay_idx = len(self.layers) - 1 - self.layerNum
        self.layer = self.layers[self.lay_idx]
        self.roundNum = 0

    # z and the round protocol -------------------------------------------------------
    def set_z(self, z1, z2, z1_2, muls, project_line: bool) -> None:
        lay = self.layer
        self.project_line = project_line
        self.w1, self.w2, self.w3 = [], [], []
        self.beta_init = nb.compute_beta(z2, 1)
        if z1_2 is not None:
            c1 = nb.compute_beta(z1, muls[0])
            c2 = nb.compute_beta(z1_2, muls[1])
            zr = np.array([(int(a) + int(b)) % P for a, b in zip(c1.real[:lay.n_gates],
                                                                   c2.real[:lay.n_gates])], np.uint64)
            zi = np.array([(int(a) + int(b)) % P for a, b in zip(c1.imag[:lay.n_gates],
                                                                   c2.imag[:lay.n_gates])], np.uint64)
        else:
            c1 = nb.compute_beta(z1, 1)
            zr = np.ascontiguousarray(c1.real[:lay.n_gates])
            zi = np.ascontiguousarray(c1.imag[:lay.n_gates])
        self.z1r, self.z1i = zr, zi
        self.type_idx = lay.type_idx(self.muxbits)
        self._start_early()

    def _start_early(self):
        torch = _torch()
        lay = self.layer
        cp = 1 << self.nCopyBits
        n_wires = 1 << lay.in_bits
        src = self.values[self.lay_idx]
        re = torch.empty((n_wires, cp), dtype=torch.int64, device="cuda")
        im = torch.empty((n_wires, cp), dtype=torch.int64, device="cuda")
        _check(nb.LIB.tp_wire_table(ctypes.c_void_p(src.data_ptr()), src.shape[1], self.nCopies,
                                    cp, n_wires, ctypes.c_void_p(re.data_ptr()),
                                    ctypes.c_void_p(im.data_ptr()), None), "tp_wire_table")
        torch.cuda.synchronize()
        self.session = None
        self.late = None
        if self.nCopyBits < 1:
            raise ValueError("nCopies must be at least 2")
        br, bi = (np.ascontiguousarray(x) for x in nb.as_coords(self.beta_init))
        h = nb.LIB.tp_early_create_dev(lay.n_gates, _p(self.type_idx), _p(lay.in0), _p(lay.in1),
                                       _p(self.z1r), _p(self.z1i), cp, n_wires,
                                       ctypes.c_void_p(re.data_ptr()), ctypes.c_void_p(im.data_ptr()),
                                       _p(br), _p(bi))
        if not h:
            raise RuntimeError("tp_early_create_dev failed: %s" % nb.last_error())
        self.session = nb.EarlySession(h, n_wires)

    def _start_late(self, wr, wi, beta):
        lay = self.layer
        self.V = (np.ascontiguousarray(wr, np.uint64), np.ascontiguousarray(wi, np.uint64))
        h = nb.LIB.tp_late_create(lay.n_gates, _p(self.type_idx), _p(lay.in0), _p(lay.in1),
                                  _p(self.z1r), _p(self.z1i), lay.in_bits, _p(self.V[0]),
                                  _p(self.V[1]), beta.real, beta.imag)
        if not h:
            raise RuntimeError("tp_late_create failed: %s" % nb.last_error())
        self.late = ctypes.c_void_p(h)
        self._pending_r = None

    def get_outputs(self):
        lay = self.layer
        nin = lay.in_bits
        total = self.nCopyBits + 2 * nin
        if self.roundNum == total:
            return self._h_output()
        if self.roundNum < self.nCopyBits:
            return interpolate_cubic(self.session.claim())
        out6 = np.zeros(6, np.uint64)
        rd = self.roundNum - self.nCopyBits
        r = self._pending_r
        rp = _fp2_arr([r]) if r is not None else None
        _check(nb.LIB.tp_late_round(self.late, rd, _p(rp) if rp is not None else None, _p(out6)),
               "tp_late_round")
        vals = [_raw(int(out6[2 * t]), int(out6[2 * t + 1])) for t in range(3)]
        return interpolate_quadratic(vals)

    def next_round(self, val) -> None:
        lay = self.layer
        nin = lay.in_bits
        assert self.roundNum < self.nCopyBits + 2 * nin
        if self.roundNum < self.nCopyBits:
            self.w3.append(val)
            self.session.fold(val)
            if self.roundNum == self.nCopyBits - 1:
                wr, wi, beta = self.session.finish_arrays()
                self.session.destroy()
                self.session = None
                self._start_late(wr, wi, beta)
        else:
            if self.roundNum < self.nCopyBits + nin:
                self.w1.append(val)
            else:
                self.w2.append(val)
            self._pending_r = val
        self.roundNum += 1

    def _values_at(self, points) -> List[Fp2]:
        nin = self.layer.in_bits
        pts = _fp2_arr([c for pt in points for c in pt])
        out = np.zeros(2 * len(points), np.uint64)
        _check(nb.LIB.tp_late_values(self.late, _p(pts), len(points), _p(out)), "tp_late_values")
        return [_raw(int(out[2 * k]), int(out[2 * k + 1])) for k in range(len(points))]

    def _h_output(self):
        nin = self.layer.in_bits
        w1, w2 = self.w1, self.w2
        if not self.project_line:
            return self._values_at([w1, w2])
        diff = [(b - a) % P for a, b in zip(w1, w2)]
        pts = [w1, w2] + [[(a + t * d) % P for a, d in zip(w1, diff)] for t in range(2, nin + 1)]
        return lagrange_coeffs(self._values_at(pts))

    def next_layer_point(self, tau):
        return [(a + (b - a) * tau) % P for a, b in zip(self.w1, self.w2)]

    def advance_layer(self) -> None:
        if self.late is not None:
            nb.LIB.tp_late_destroy(self.late)
            self.late = None
        self.layerNum += 1
        self._begin_layer()

    def close(self):
        if getattr(self, "late", None) is not None:
            nb.LIB.tp_late_destroy(self.late)
            self.late = None
        if getattr(self, "session", None) is not None:
            self.session.destroy()
            self.session = None


def mle_device(rows_tensor, point) -> Fp2:
    """MLE of a device-resident base-field vector (rows, copy-major) at point."""
    z = _fp2_arr(point)
    out = np.zeros(2, np.uint64)
    _check(nb.LIB.tp_mle_eval_base_dev(ctypes.c_void_p(rows_tensor.data_ptr()), len(point), _p(z),
                                       _p(out), None), "tp_mle_eval_base_dev")
    return _raw(int(out[0]), int(out[1]))


def wiring_mlext(layer: Layer, type_idx, z1, z1_2, muls, w1, w2):
    out = np.zeros(26, np.uint64)
    use2 = z1_2 is not None
    m0 = _fp2_arr([muls[0]]) if use2 else None
    m1 = _fp2_arr([muls[1]]) if use2 else None
    _check(nb.LIB.tp_wiring_mlext(layer.n_gates, _p(type_idx), _p(layer.in0), _p(layer.in1),
                                  _p(_fp2_arr(z1)), _p(_fp2_arr(z1_2)) if use2 else None,
                                  len(z1), _p(_fp2_arr(w1)), _p(_fp2_arr(w2)), len(w1),
                                  _p(m0) if use2 else None, _p(m1) if use2 else None,
                                  1 if use2 else 0, _p(out)), "tp_wiring_mlext")
    return [_raw(int(out[2 * t]), int(out[2 * t + 1])) for t in range(13)]


def mlx_z2_factor(w3, z2):
    acc = 1
    for w, z in zip(w3, z2):
        acc = acc * (2 * w * z + 1 - (w + z)) % P
    return acc


def refuse_mux(layers) -> None:
    """The NIZK provers take no MUX gates.  The reference's cannot prove one:
    its CircuitProverNIZK.run reaches eval_mlext, which reads per-layer mux
    bits (TP/libTruePix/circuitverifier.py:276) that the prover never sets, and
    raises; no proof of a MUX circuit exists to match."""
    for lay in layers:
        if any(o == "MUX" for o in lay.ops):
            raise ValueError("MUX gates cannot be proved by the NIZK prover (the reference's "
                             "CircuitProverNIZK.run fails on them as well)")


class _FromPWS(type):
    def from_pws(cls, parsed, nCopies):
        input_layer, layers = parse_layers(parsed)
        return input_layer, cls(nCopies, len(input_layer), layers)


class _NIZKBase(metaclass=_FromPWS):
    def __init__(self, nCopies, nInputs, layers: List[Layer]):
        self.nCopies = nCopies
        self.nCopyBits = clog2(nCopies)
        self.nInputs = nInputs
        self.nInBits = clog2(nInputs)
        self.layers = layers
        self.layOutBits = [l.out_bits for l in reversed(layers)]
        self.layInBits = self.layOutBits[1:] + [self.nInBits]
        self.com = ExtVOLECommit(P, None, security_repetitions=1)
        self.fs = None
        self.muxbits = None


class CircuitProverNIZK(_NIZKBase):
    """Prover of TP/libTruePix/circuitnizkvec_vole.py:21-501 on the device engine."""

    def __init__(self, nCopies, nInputs, layers):
        super().__init__(nCopies, nInputs, layers)
        self.fs = FiatShamir(P, protocol="truepix")
        self.v_elem, self.out_elem = [], []
        self.prover = None
        self.pid = None                 # transcript pid (defaults to os.getpid())
        self.vole_dir = "."
        self.manifest_path = None
        self.final_input_point = None
        self.final_input_tag = None
        self.final_input_value = None

    def build_prover(self):
        self.prover = CircuitProver(self.nCopies, self.nInputs, self.layers)

    def _challenge(self, label):
        return self.fs.rand_ext(label)

    def create_pok_temp(self, outs):
        rvals = []
        for val in outs:
            v_m, d_m = self.com.commit(val)
            rvals.append(v_m)
            self.fs.put(d_m)
            if len(self.v_elem) < len(outs):
                self.v_elem.append(0)
                self.out_elem.append(0)
        for k in range(len(rvals)):
            self.v_elem[k] += rvals[k]
            self.out_elem[k] += outs[k]
        return rvals

    def create_pok_final(self):
        chal = self._challenge("pok-final")
        z1, z2 = self.com.pok_finish(sum(self.out_elem), sum(self.v_elem), chal)
        self.fs.put((z1, z2))

    def create_prod_pok(self, outs):
        prod = (outs[0] * outs[1]) % P
        v1, d1 = self.com.commit(outs[0])
        v2, d2 = self.com.commit(outs[1])
        v3, d3 = self.com.commit(prod)
        self.fs.put((d1, d2, d3))
        self.fs.put(self.com.prod_finish((outs[0], outs[1], prod), (v1, v2, v3)))
        return (v1, v2, v3)

    def create_final_prod_pok(self, outs):
        v1, d1 = self.com.commit(outs[0])
        sval = sum(outs) % P
        prod = (outs[0] * sval) % P
        v3, d3 = self.com.commit(prod)
        self.fs.put((d1, d3))
        rest_v = []
        for val in outs[1:]:
            v_m, d_m = self.com.commit(val)
            rest_v.append(v_m)
            self.fs.put(d_m)
        chal = self._challenge("pok-vec")
        z1, z2 = self.com.pok_finish(sum(outs[1:]) % P, sum(rest_v) % P, chal)
        self.fs.put((z1, z2))
        vvals = [v1] + rest_v
        pr = (v1, sum(vvals) % P, v3)
        self.fs.put(self.com.prod_finish((outs[0], sval, prod), pr))
        return vvals, pr

    def create_eq_proof(self, v1val, v2val):
        chal = self._challenge("eq")
        v_z = self.com.est_val(v2val, chal) if v1val is None else self.com.est_eq(v1val, v2val, chal)
        self.fs.put(v_z)

    def create_input_value_proof(self, tag, value):
        chal = self._challenge("input-val")
        self.fs.put(self.com.est_val(tag, chal))

    def run(self, inputs, muxbits=None):
        refuse_mux(self.layers)
        import time as _time
        marks = [("start", _time.perf_counter())]

        def mark(name):
            marks.append((name, _time.perf_counter()))
        if self.prover is None:
            self.build_prover()
        pid = os.getpid() if self.pid is None else self.pid
        self.fs.put(pid)
        self.com.load_from_file_prover(os.path.join(self.vole_dir, "vole_ext_prover_%d.json" % pid))
        prover = self.prover
        if muxbits is not None:
            prover.muxbits = list(muxbits)
        self.fs.put(muxbits, True)
        n_inv = 1 << (self.nInBits + self.nCopyBits)
        self.fs.put(["private_input_omitted", n_inv], True)
        mark("setup")
        prover.set_inputs(inputs)
        mark("circuit_eval")
        outvals = prover.ckt_outputs.reshape(-1)
        self.fs.put_io_int_array(outvals)
        self.fs.put(commit_binding(read_commit_manifest(self.manifest_path)))
        mark("absorb_outputs")

        nOutBits = self.layOutBits[0]
        z1 = [self._challenge("z1-%d" % i) for i in range(nOutBits)]
        z1_2, muls = None, None
        z2 = [self._challenge("z2-%d" % i) for i in range(self.nCopyBits)]
        project_line = len(self.layers) == 1
        prover.set_z(z1, z2, None, None, project_line)
        mark("tables")
        prev_rval = None
        for lay in range(len(self.layers)):
            nInBits = self.layInBits[lay]
            w1, w2, w3 = [], [], []
            for rd in range(2 * nInBits + self.nCopyBits):
                outs = prover.get_outputs()
                outs_rvals = self.create_pok_temp(outs)
                zp1 = (sum(outs_rvals) + outs_rvals[0]) % P
                self.create_eq_proof(prev_rval, zp1)
                nrand = self._challenge("sc-%d-%d" % (lay, rd))
                prover.next_round(nrand)
                prev_rval = horner_eval(outs_rvals, nrand)
                if rd < self.nCopyBits:
                    w3.append(nrand)
                elif rd < self.nCopyBits + nInBits:
                    w1.append(nrand)
                else:
                    w2.append(nrand)
            outs = prover.get_outputs()
            if project_line:
                outs_rvals, pr_rvals = self.create_final_prod_pok(outs)
            else:
                pr_rvals = self.create_prod_pok(outs)
            layer = prover.layer
            mlext = wiring_mlext(layer, prover.type_idx, z1, z1_2, muls, w1, w2)
            mlx = mlx_z2_factor(w3, z2)
            tv = 0
            for idx, elm in enumerate(mlext):
                tv = (tv + elm * GATE_COM_P[idx](pr_rvals[0], pr_rvals[1], pr_rvals[2])) % P
            tv = (tv * mlx) % P
            self.create_eq_proof(prev_rval, tv)
            project_next = lay == len(self.layers) - 2
            if project_line:
                tau = self._challenge("tau-%d" % lay)
                muls = None
                prev_rval = horner_eval(outs_rvals, tau)
                z1 = [(a + (b - a) * tau) % P for a, b in zip(w1, w2)]
                z1_2 = None
            else:
                muls = [self._challenge("mul0-%d" % lay), self._challenge("mul1-%d" % lay)]
                prev_rval = (muls[0] * pr_rvals[0] + muls[1] * pr_rvals[1]) % P
                prover.advance_layer()
                if project_next:
                    prover.set_z(w1, w3, w2, muls, project_next)
                else:
                    prover.set_z(w1, w3, w2, muls, project_next)
                z1 = w1
                z1_2 = w2
            project_line = project_next
            z2 = w3
        mark("sumcheck")
        self.create_pok_final()
        self.final_input_point = list(z1) + list(z2)
        self.final_input_tag = prev_rval
        self.fs.put(["gkr_input_point", serialize_point(self.final_input_point)])
        self.final_input_value = mle_device(prover.input_vector_device(), self.final_input_point)
        self.create_input_value_proof(prev_rval, self.final_input_value)
        prover.close()
        mark("input_mle")
        out = self.fs.to_string()
        mark("serialize")
        self.timings = {marks[i][0]: marks[i][1] - marks[i - 1][1] for i in range(1, len(marks))}
        return out


class CircuitVerifierNIZK(_NIZKBase):
    """Verifier of TP/libTruePix/circuitnizkvec_vole.py:503-968 (GKR part),
    with the output MLE and wiring predicate on the device."""

    def __init__(self, nCopies, nInputs, layers):
        super().__init__(nCopies, nInputs, layers)
        self.p_elem = []
        self.pending_input_claim = None
        self.vole_dir = "."
        self.output_path = "output.txt"
        self.manifest_path = None
        self.mlx_w1 = self.mlx_w2 = None

    def _challenge(self, label):
        return self.fs.rand_ext(label)

    def check_pok_temp(self, n):
        pvals = []
        for i in range(n):
            pvals.append(self.com.change_commit(self.fs.take()[0]))
            if len(self.p_elem) < n:
                self.p_elem.append(0)
            self.p_elem[i] += pvals[i]
        return pvals

    def check_pok_final(self):
        chal = self._challenge("pok-final")
        z1, z2 = self.fs.take()[0]
        return self.com.pok_check(z1, z2, chal, sum(self.p_elem))

    def check_prod_pok(self):
        pvals = [self.com.change_commit(d) for d in self.fs.take()[0]]
        avals = self.fs.take()[0]
        return pvals, self.com.prod_check(pvals, avals)

    def check_final_prod_pok(self, nInBits):
        d1, d3 = self.fs.take()[0]
        p1, p3 = self.com.change_commit(d1), self.com.change_commit(d3)
        rest = [self.com.change_commit(self.fs.take()[0]) for _ in range(nInBits)]
        chal = self._challenge("pok-vec")
        z1, z2 = self.fs.take()[0]
        ok = self.com.pok_check(z1, z2, chal, sum(rest) % P)
        pvals = [p1] + rest
        p2 = sum(pvals) % P
        avals = self.fs.take()[0]
        ok = ok and self.com.prod_check((p1, p2, p3), avals)
        return pvals, p2, p3, ok

    def check_val_proof(self, qval, val):
        chal = self._challenge("eq")
        return self.com.est_val_check(qval, chal, val, self.fs.take()[0])

    def check_eq_proof(self, q1, q2):
        chal = self._challenge("eq")
        return self.com.est_eq_check(q1, q2, chal, self.fs.take()[0])

    def read_outputs(self):
        """output.txt: 384 values per line, each row padded with 128 zeros
        (circuitnizkvec_vole.py:714-755)."""
        rows = []
        with open(self.output_path) as fh:
            for ln, line in enumerate(fh, 1):
                s = line.strip()
                if not s:
                    continue
                data = s.split()
                if len(data) != 384:
                    raise ValueError("Line %d: Expected 384 numbers, found %d" % (ln, len(data)))
                rows.append([int(x) for x in data] + [0] * 128)
        vec = np.array([v for r in rows for v in r], dtype=np.uint64) if rows else np.zeros(0, np.uint64)
        n = 1 << clog2(len(vec)) if len(vec) else 1
        out = np.zeros(n, dtype=np.uint64)
        out[:len(vec)] = vec
        return out

    def run(self, pf):
        torch = _torch()
        self.fs = FiatShamir.from_string(pf, protocol="truepix")
        pid = int(self.fs.take()[0])
        self.com.load_from_file_verifier(os.path.join(self.vole_dir, "vole_ext_verifier_%d.json" % pid))
        self.muxbits = self.fs.take(True)
        self.fs.take(True)                     # input metadata
        outputs = self.read_outputs()
        self.fs.take(True)                     # the prover's outputs (hashed)
        claimed = self.fs.take()
        if list(claimed) != commit_binding(read_commit_manifest(self.manifest_path)):
            raise ValueError("Orion commitment in the proof does not match the commitment manifest")
        nOutBits = self.layOutBits[0]
        if clog2(len(outputs)) != nOutBits + self.nCopyBits:
            raise ValueError("output.txt has the wrong number of values")
        z1 = [self._challenge("z1-%d" % i) for i in range(nOutBits)]
        z1_2, muls = None, None
        z2 = [self._challenge("z2-%d" % i) for i in range(self.nCopyBits)]
        project_line = len(self.layers) == 1
        dev_out = torch.from_numpy(outputs.view(np.int64)).to("cuda")
        expect = mle_device(dev_out, z1 + z2)
        prev = None
        for lay in range(len(self.layers)):
            nInBits = self.layInBits[lay]
            w1, w2, w3 = [], [], []
            for rd in range(2 * nInBits + self.nCopyBits):
                cvals = self.check_pok_temp(4 if rd < self.nCopyBits else 3)
                ncom = (sum(cvals) + cvals[0]) % P
                ok = self.check_val_proof(ncom, expect) if prev is None else self.check_eq_proof(prev, ncom)
                if not ok:
                    raise ValueError("Verification failed in round %d of layer %d" % (rd, lay))
                nrand = self._challenge("sc-%d-%d" % (lay, rd))
                prev = horner_eval(cvals, nrand)
                (w3 if rd < self.nCopyBits else (w1 if rd < self.nCopyBits + nInBits else w2)).append(nrand)
            if project_line:
                cvals, c2, c3, ok = self.check_final_prod_pok(nInBits)
                if not ok:
                    raise ValueError("Verification of final product PoK failed")
                pr = (cvals[0], c2, c3)
            else:
                pr, ok = self.check_prod_pok()
                if not ok:
                    raise ValueError("Verification of product PoK failed in layer %d" % lay)
            layer = self.layers[len(self.layers) - 1 - lay]
            tidx = layer.type_idx(self.muxbits if self.muxbits is not None else [0] * 64)
            mlext = wiring_mlext(layer, tidx, z1, z1_2, muls, w1, w2)
            tv = self.com.tV_eval(pr, mlext, mlx_z2_factor(w3, z2))
            if not self.check_eq_proof(prev, tv):
                raise ValueError("Verification of mlext eq proof failed in layer %d" % lay)
            project_next = lay == len(self.layers) - 2
            if project_line:
                tau = self._challenge("tau-%d" % lay)
                muls = None
                prev = horner_eval(cvals, tau)
                z1 = [(a + (b - a) * tau) % P for a, b in zip(w1, w2)]
                z1_2 = None
            else:
                muls = [self._challenge("mul0-%d" % lay), self._challenge("mul1-%d" % lay)]
                prev = (muls[0] * pr[0] + muls[1] * pr[1]) % P
                z1, z1_2 = w1, w2
            project_line = project_next
            z2 = w3
        if not self.check_pok_final():
            raise ValueError("PoK failed for commits ")
        self.pending_input_claim = PendingInputClaim(input_point=list(z1) + list(z2),
                                                     authenticated_claim=prev,
                                                     n_copies=self.nCopies, n_in_bits=self.nInBits)
        self.fs.take()
        return self.pending_input_claim

    def finish_input(self, value):
        if self.pending_input_claim is None:
            raise RuntimeError("finish_input called without pending claim")
        chal = self._challenge("input-val")
        if not self.com.est_val_check(self.pending_input_claim.authenticated_claim, chal, value,
                                      self.fs.take()[0]):
            raise ValueError("Input opening check against Orion value failed")
        return True


ProverClass = CircuitProverNIZK
VerifierClass = CircuitVerifierNIZK
