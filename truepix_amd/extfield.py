"""Fp2 = F_p[i] / (i^2 + 1), p = 2^61 - 1: the host-side field element.

Interface of TP/libTruePix/extfield.py:29-196 (constructors, arithmetic with
ints and Fp2, inverse, serialize) so host protocol code reads the same.  The
class pickles under the name ``libTruePix.extfield.Fp2`` with state
``(real, imag, p)`` because that is the name and state the reference's proof
format carries (TP/libTruePix/extfield.py:76-82) and its verifier unpickles.
"""
from __future__ import annotations

import secrets
import sys
import types
from typing import Iterable, List, Sequence, Tuple

BASE_PRIME_61 = (1 << 61) - 1
P = BASE_PRIME_61


class Fp2:
    __slots__ = ("real", "imag", "p")

    def __init__(self, real: int, imag: int, p: int = P):
        self.real = real % p
        self.imag = imag % p
        self.p = p

    # constructors ------------------------------------------------------------
    @classmethod
    def zero(cls, p: int = P) -> "Fp2":
        return _raw(0, 0, p)

    @classmethod
    def one(cls, p: int = P) -> "Fp2":
        return _raw(1, 0, p)

    @classmethod
    def from_base(cls, x: int, p: int = P) -> "Fp2":
        return _raw(int(x) % p, 0, p)

    @classmethod
    def random(cls, p: int = P, nonzero: bool = False) -> "Fp2":
        while True:
            re, im = secrets.randbelow(p), secrets.randbelow(p)
            if not nonzero or re or im:
                return _raw(re, im, p)

    # inspection ----------------------------------------------------------------
    def is_base(self) -> bool:
        return self.imag == 0

    def to_pair(self) -> Tuple[int, int]:
        return (self.real, self.imag)

    def serialize(self) -> List[int]:
        return [self.real, self.imag]

    @classmethod
    def deserialize(cls, data: Sequence[int], p: int = P) -> "Fp2":
        if len(data) != 2:
            raise ValueError("a serialized Fp2 is a pair of ints, got %d items" % len(data))
        return cls(data[0], data[1], p)

    # pickling: state (real, imag, p), as in the proof format
    def __getstate__(self):
        return (self.real, self.imag, self.p)

    def __setstate__(self, state):
        self.real, self.imag, self.p = state

    # comparisons ---------------------------------------------------------------
    def __eq__(self, other) -> bool:
        if type(other) is Fp2:
            return self.real == other.real and self.imag == other.imag and self.p == other.p
        if isinstance(other, int):
            return self.imag == 0 and self.real == other % self.p
        return NotImplemented

    def __hash__(self) -> int:
        return hash((self.real, self.imag))

    # arithmetic ----------------------------------------------------------------
    def __neg__(self) -> "Fp2":
        p = self.p
        return _raw(-self.real % p, -self.imag % p, p)

    def __add__(self, o) -> "Fp2":
        p = self.p
        if type(o) is Fp2:
            return _raw((self.real + o.real) % p, (self.imag + o.imag) % p, p)
        if isinstance(o, int):
            # adding the integer 0 returns the same object, as the reference does
            # (sum() over Fp2 therefore starts from the first element itself)
            if o == 0:
                return self
            return _raw((self.real + o) % p, self.imag, p)
        return NotImplemented

    __radd__ = __add__

    def __sub__(self, o) -> "Fp2":
        p = self.p
        if type(o) is Fp2:
            return _raw((self.real - o.real) % p, (self.imag - o.imag) % p, p)
        if isinstance(o, int):
            return _raw((self.real - o) % p, self.imag, p)
        return NotImplemented

    def __rsub__(self, o) -> "Fp2":
        p = self.p
        return _raw((o - self.real) % p, -self.imag % p, p)

    def __mul__(self, o) -> "Fp2":
        p = self.p
        if type(o) is Fp2:
            a, b, c, d = self.real, self.imag, o.real, o.imag
            return _raw((a * c - b * d) % p, (a * d + b * c) % p, p)
        if isinstance(o, int):
            return _raw(self.real * o % p, self.imag * o % p, p)
        return NotImplemented

    __rmul__ = __mul__

    def __truediv__(self, o) -> "Fp2":
        return self * _as_fp2(o, self.p).inverse()

    def __mod__(self, q: int) -> "Fp2":
        if q != self.p:
            raise ValueError("reducing an Fp2 element modulo %d: only its own prime is allowed" % q)
        return self

    def inverse(self) -> "Fp2":
        p = self.p
        n = (self.real * self.real + self.imag * self.imag) % p
        if n == 0:
            raise ZeroDivisionError("0 has no multiplicative inverse")
        ni = pow(n, p - 2, p)
        return _raw(self.real * ni % p, -self.imag * ni % p, p)

    def __repr__(self) -> str:
        return "Fp2(%d, %d)" % (self.real, self.imag)


_NEW = Fp2.__new__


def _raw(real: int, imag: int, p: int = P) -> Fp2:
    v = _NEW(Fp2)
    v.real = real
    v.imag = imag
    v.p = p
    return v


def _as_fp2(v, p: int = P) -> Fp2:
    if type(v) is Fp2:
        return v
    if isinstance(v, int):
        return _raw(v % p, 0, p)
    raise TypeError("cannot convert %r to Fp2" % (type(v),))


def fold_val(in0, in1, val, p: int = P) -> Fp2:
    """The multilinear fold of the pair (in0, in1) at val: the line through
    in0 (at 0) and in1 (at 1) evaluated at val.  Operands are ints or Fp2;
    the result is always a new Fp2."""
    lo, hi, x = _as_fp2(in0, p), _as_fp2(in1, p), _as_fp2(val, p)
    return _raw(lo.real, lo.imag, p) + (hi - lo) * x


def serialize_point(point: Sequence[Fp2]) -> List[List[int]]:
    return [e.serialize() for e in point]


def deserialize_point(data: Sequence[Sequence[int]], p: int = P) -> List[Fp2]:
    return [Fp2.deserialize(x, p) for x in data]


def embed_base_list(values: Iterable[int], p: int = P) -> List[Fp2]:
    return [Fp2.from_base(v, p) for v in values]


# --- the proof format's class name -------------------------------------------
PICKLE_MODULE = "libTruePix.extfield"
Fp2.__module__ = PICKLE_MODULE


class proof_pickle_names:
    """Context manager: make ``libTruePix.extfield.Fp2`` resolve to this class
    while a proof is pickled (pickle looks the class up by that name).  A real
    ``libTruePix`` already imported is left untouched."""

    def __enter__(self):
        self._added = []
        if PICKLE_MODULE in sys.modules:
            mod = sys.modules[PICKLE_MODULE]
            if getattr(mod, "Fp2", None) is not Fp2:
                raise RuntimeError("a different libTruePix.extfield is imported")
            return self
        pkg = sys.modules.get("libTruePix")
        if pkg is None:
            pkg = types.ModuleType("libTruePix")
            pkg.__path__ = []
            sys.modules["libTruePix"] = pkg
            self._added.append("libTruePix")
        mod = types.ModuleType(PICKLE_MODULE)
        mod.Fp2 = Fp2
        sys.modules[PICKLE_MODULE] = mod
        self._added.append(PICKLE_MODULE)
        return self

    def __exit__(self, *exc):
        for name in reversed(self._added):
            sys.modules.pop(name, None)
        return False
