"""Extension-field VOLE commitments (prover and verifier views) and the EVOL
correlation file format.

Same operations and file layout as TP/libTruePix/ExtVOLECommit.py:23-338 and
TP/libTruePix/ExtVOLEGen.py:23-187, so proofs and files interoperate with the
reference:
  correlation  P = U * Delta + V in Fp2 (prover holds U, V; verifier Delta, P)
  commit(m)    consumes the next correlation: tag = V, correction d = m - U
  EVOL file    header <4sIIIIQ> = ("EVOL", version 1, role 1|2, reps, n, p),
               prover: U.re U.im (per rep), V.re V.im (per rep) as u64 arrays;
               verifier: deltas (re, im per rep), then P.re P.im (per rep).
The generator here can draw from the OS CSPRNG (as the reference does) or
from a seeded PCG64 stream for reproducible fixtures and benchmarks.
"""
from __future__ import annotations

import secrets
import struct
from typing import List, Optional

import numpy as np

from truepix_amd.extfield import BASE_PRIME_61, Fp2, _as_fp2, _raw

P = BASE_PRIME_61
MAGIC = b"EVOL"
VERSION = 1
ROLE_PROVER = 1
ROLE_VERIFIER = 2
HEADER = struct.Struct("<4sIIIIQ")
DEFAULT_VOLE_INSTANCES = 500


def _mulmod_arrays(ar, ai, br, bi):
    """(ar + ai i)(br + bi i) mod p for uint64 arrays, via Python ints (setup only)."""
    out_r = np.empty(ar.size, dtype=np.uint64)
    out_i = np.empty(ar.size, dtype=np.uint64)
    for k, (a, b, c, d) in enumerate(zip(ar.tolist(), ai.tolist(), br.tolist(), bi.tolist())):
        out_r[k] = (a * c - b * d) % P
        out_i[k] = (a * d + b * c) % P
    return out_r, out_i


class ExtVOLEGenerator:
    """Generates Delta and n correlations per security repetition."""

    def __init__(self, base_prime: int = P, security_repetitions: int = 1,
                 seed: Optional[int] = None):
        if base_prime != P:
            raise ValueError("only p = 2^61 - 1 is supported")
        self.security_repetitions = max(1, int(security_repetitions))
        self._rng = np.random.Generator(np.random.PCG64(seed)) if seed is not None else None
        self.deltas: List[Fp2] = []
        self.u = self.v = self.p = None   # lists over reps of (re, im) arrays

    def _nonzero_pairs(self, n: int):
        re = np.empty(n, dtype=np.uint64)
        im = np.empty(n, dtype=np.uint64)
        for k in range(n):
            while True:
                if self._rng is None:
                    a, b = secrets.randbelow(P), secrets.randbelow(P)
                else:
                    a, b = (int(x) for x in self._rng.integers(0, P, size=2, dtype=np.int64))
                if a or b:
                    break
            re[k], im[k] = a, b
        return re, im

    def generate_deltas(self) -> List[Fp2]:
        re, im = self._nonzero_pairs(self.security_repetitions)
        self.deltas = [_raw(int(a), int(b)) for a, b in zip(re, im)]
        return list(self.deltas)

    def generate_vole_instances(self, num_instances: int):
        if not self.deltas:
            raise ValueError("Call generate_deltas() first")
        self.u, self.v, self.p = [], [], []
        for d in self.deltas:
            ur, ui = self._nonzero_pairs(num_instances)
            vr, vi = self._nonzero_pairs(num_instances)
            dr = np.full(num_instances, d.real, dtype=np.uint64)
            di = np.full(num_instances, d.imag, dtype=np.uint64)
            tr, ti = _mulmod_arrays(ur, ui, dr, di)
            pr = ((tr.astype(object) + vr.astype(object)) % P).astype(np.uint64)
            pi = ((ti.astype(object) + vi.astype(object)) % P).astype(np.uint64)
            self.u.append((ur, ui))
            self.v.append((vr, vi))
            self.p.append((pr, pi))
        return (len(self.u), num_instances)

    def save_to_files(self, prover_file: str, verifier_file: str, session_id: str = ""):
        del session_id
        if not self.deltas or not self.u:
            raise ValueError("No VOLE instances generated yet")
        n = self.u[0][0].size
        reps = self.security_repetitions
        with open(prover_file, "wb") as fh:
            fh.write(HEADER.pack(MAGIC, VERSION, ROLE_PROVER, reps, n, P))
            for re, im in self.u + self.v:
                fh.write(re.astype("<u8").tobytes())
                fh.write(im.astype("<u8").tobytes())
        with open(verifier_file, "wb") as fh:
            fh.write(HEADER.pack(MAGIC, VERSION, ROLE_VERIFIER, reps, n, P))
            fh.write(np.array([c for d in self.deltas for c in (d.real, d.imag)],
                              dtype="<u8").tobytes())
            for re, im in self.p:
                fh.write(re.astype("<u8").tobytes())
                fh.write(im.astype("<u8").tobytes())


def read_vole_file(filename: str):
    """Parse an EVOL file (TP/libTruePix/ExtVOLEGen.py:156-187)."""
    with open(filename, "rb") as fh:
        head = fh.read(HEADER.size)
        if head[:4] != MAGIC:
            raise ValueError("%s is not an EVOL correlation file" % filename)
        if len(head) != HEADER.size:
            raise ValueError("truncated VOLE header")
        _, version, role, reps, n, base = HEADER.unpack(head)
        if version != VERSION:
            raise ValueError("unsupported VOLE file version %s" % version)
        if base != P:
            raise ValueError("VOLE file is for another prime")

        def arr(k):
            raw = fh.read(8 * k)
            if len(raw) != 8 * k:
                raise ValueError("truncated VOLE file")
            return np.frombuffer(raw, dtype="<u8").astype(np.uint64)
        deltas = []
        if role == ROLE_VERIFIER:
            d = arr(2 * reps)
            deltas = [(int(d[2 * i]), int(d[2 * i + 1])) for i in range(reps)]
            nstreams = 1
        elif role == ROLE_PROVER:
            nstreams = 2
        else:
            raise ValueError("unknown VOLE role %s" % role)
        streams = [(arr(n), arr(n)) for _ in range(reps * nstreams)]
    return {"role": role, "security_repetitions": reps, "n": n, "deltas": deltas,
            "streams": streams}


class ExtVOLECommit:
    """VOLE-MAC commitments over Fp2 (TP/libTruePix/ExtVOLECommit.py:23-338)."""

    def __init__(self, base_prime: int = P, rec=None, fs=None, security_repetitions: int = 1):
        self.p = self.q = int(base_prime)
        self.rec = rec
        self.security_repetitions = max(1, int(security_repetitions))
        self.deltas: List[Fp2] = []
        self.u_lists = []
        self.v_lists = []
        self.p_lists = []
        self.numVOLE = 0
        self.VOLEindex = 0

    # loading ------------------------------------------------------------------
    @staticmethod
    def _to_lists(streams):
        return [[_raw(int(a), int(b)) for a, b in zip(re.tolist(), im.tolist())]
                for re, im in streams]

    def load_from_file_prover(self, filename: str) -> int:
        blob = read_vole_file(filename)
        if blob["role"] != ROLE_PROVER:
            raise ValueError("Not an ExtVOLE prover file for truepix")
        reps = self.security_repetitions = int(blob["security_repetitions"])
        self.u_lists = self._to_lists(blob["streams"][:reps])
        self.v_lists = self._to_lists(blob["streams"][reps:])
        self.VOLEindex = blob["n"] if self.u_lists else 0
        self.numVOLE = 0
        return self.VOLEindex

    def load_from_file_verifier(self, filename: str) -> int:
        blob = read_vole_file(filename)
        if blob["role"] != ROLE_VERIFIER:
            raise ValueError("Not an ExtVOLE verifier file for truepix")
        self.security_repetitions = int(blob["security_repetitions"])
        self.deltas = [_raw(r, i) for r, i in blob["deltas"]]
        self.p_lists = self._to_lists(blob["streams"])
        self.VOLEindex = blob["n"] if self.p_lists else 0
        self.numVOLE = 0
        return self.VOLEindex

    def _pack(self, values):
        return values[0] if self.security_repetitions == 1 else values

    def _unpack(self, value) -> List[Fp2]:
        if isinstance(value, list):
            return [_as_fp2(v, self.p) for v in value]
        return [_as_fp2(value, self.p)] * self.security_repetitions

    def _consume_index(self) -> int:
        if self.VOLEindex <= 0:
            raise IndexError("VOLE instance index out of range")
        idx = self.numVOLE
        self.numVOLE += 1
        self.VOLEindex -= 1
        return idx

    def used_vole_stats(self):
        count = self.numVOLE * self.security_repetitions
        return count, count * (32 if self.u_lists else 16)

    # prover -----------------------------------------------------------------------
    def commit(self, m):
        msg = _as_fp2(m, self.p)
        idx = self._consume_index()
        tags = [self.v_lists[r][idx] for r in range(self.security_repetitions)]
        ds = [msg - self.u_lists[r][idx] for r in range(self.security_repetitions)]
        return self._pack(tags), self._pack(ds)

    def pok_finish(self, m, v_m, c):
        msg, chal = _as_fp2(m, self.p), _as_fp2(c, self.p)
        return msg * chal, self._pack([t * chal for t in self._unpack(v_m)])

    def est_eq(self, v1, v2, c):
        chal = _as_fp2(c, self.p)
        t1, t2 = self._unpack(v1), self._unpack(v2)
        return self._pack([(t1[r] - t2[r]) * chal for r in range(self.security_repetitions)])

    def est_val(self, v_r, c):
        chal = _as_fp2(c, self.p)
        return self._pack([t * chal for t in self._unpack(v_r)])

    def prod_finish(self, xvals, vval):
        x1, x2, _ = [_as_fp2(x, self.p) for x in xvals]
        v1, v2, v3 = (self._unpack(v) for v in vval)
        a1 = [v1[r] * v2[r] for r in range(self.security_repetitions)]
        a2 = [x1 * v2[r] + x2 * v1[r] - v3[r] for r in range(self.security_repetitions)]
        return (self._pack(a1), self._pack(a2))

    # verifier -----------------------------------------------------------------
    def change_commit(self, d_m):
        ds = self._unpack(d_m)
        idx = self._consume_index()
        return self._pack([self.p_lists[r][idx] + ds[r] * self.deltas[r]
                           for r in range(self.security_repetitions)])

    def pok_check(self, z1, z2, c, pval) -> bool:
        chal, z1v = _as_fp2(c, self.p), _as_fp2(z1, self.p)
        z2s, pv = self._unpack(z2), self._unpack(pval)
        return all(z1v * self.deltas[r] + z2s[r] == pv[r] * chal
                   for r in range(self.security_repetitions))

    def est_eq_check(self, q1, q2, c, v_z) -> bool:
        chal = _as_fp2(c, self.p)
        a, b, z = self._unpack(q1), self._unpack(q2), self._unpack(v_z)
        return all((a[r] - b[r]) * chal == z[r] for r in range(self.security_repetitions))

    def est_val_check(self, p_x, c, val, v_z) -> bool:
        chal, msg = _as_fp2(c, self.p), _as_fp2(val, self.p)
        pv, zs = self._unpack(p_x), self._unpack(v_z)
        return all(zs[r] == (pv[r] - msg * self.deltas[r]) * chal
                   for r in range(self.security_repetitions))

    def prod_check(self, pval, aval) -> bool:
        p1, p2, p3 = (self._unpack(v) for v in pval)
        a1, a2 = self._unpack(aval[0]), self._unpack(aval[1])
        return all(p1[r] * p2[r] - p3[r] * self.deltas[r] == a1[r] + a2[r] * self.deltas[r]
                   for r in range(self.security_repetitions))

    def tV_eval(self, cvals, mlext_evals, mlx_z2):
        from truepix_amd.gkr import GATE_COM_V
        c0, c1, c2 = (self._unpack(v) for v in cvals)
        mlx = _as_fp2(mlx_z2, self.p)
        outs = []
        for r in range(self.security_repetitions):
            acc = Fp2.zero(self.p)
            for idx, elm in enumerate(mlext_evals):
                base = GATE_COM_V[idx](c0[r], c1[r], c2[r])
                acc = acc + _as_fp2(base, self.p) * (_as_fp2(elm, self.p) * mlx)
            outs.append(acc)
        return self._pack(outs)
