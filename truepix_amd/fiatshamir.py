"""The TruePix Fiat-Shamir transcript (TP/libTruePix/fiatshamir.py:16-237).

Same interface (``put``, ``take``, ``rand_scalar``, ``rand_ext``,
``to_string``, ``from_string``, ``proof_size``) and the same absorbed bytes,
so challenges and proof bytes match the reference:

  prefix   b"TruePix," + protocol + b"," + b"binhash1,"
  list of ints (first element an int, all 0 <= x < 2^64)
           0x01 <u32 n> <u64 x n>
  otherwise each flattened element:
           int 0x02 <u64 (x mod 2^64)> | Fp2 0x03 <u64 re><u64 im>
           str 0x04 <u32 len><utf-8> | None 0x05
  rand_scalar: absorb [counter], counter += 1, take the big-endian digest's
           low 61 bits, retry while >= p (q = p needs no multiple, :38-51).
  rand_ext(label): "ext", label, "real" -> real; "ext", label, "imag" -> imag.

The proof is ``pickle.dumps((q, io, trans, ndb, rvstart, rvend, wDiv, 4), -1)``
(:234-237) with Fp2 pickled under ``libTruePix.extfield.Fp2``; loading uses a
restricted unpickler that only admits that class.
"""
from __future__ import annotations

import bz2
import hashlib
import io as _io
import pickle
import struct
from collections import deque

from truepix_amd.extfield import Fp2, _raw, proof_pickle_names

P = (1 << 61) - 1
_PACK_Q = struct.Struct("<Q").pack
_PACK_I = struct.Struct("<I").pack
_MASK64 = (1 << 64) - 1


def flatten_iter(vals):
    """Depth-first flattening of nested lists / tuples (and Fp2Buf)."""
    from truepix_amd.native_fp2_bridge import Fp2Buf
    for v in vals:
        if isinstance(v, (list, tuple, Fp2Buf)):
            yield from flatten_iter(v)
        else:
            yield v


def encode_items(vals) -> bytes:
    """Bytes one hash_update(vals) absorbs (binary mode)."""
    if isinstance(vals, list) and vals and type(vals[0]) is int:
        if all(type(v) is int and 0 <= v <= _MASK64 for v in vals):
            return b"\x01" + _PACK_I(len(vals)) + struct.pack("<%dQ" % len(vals), *vals)
    out = bytearray()
    for e in flatten_iter(vals):
        t = type(e)
        if t is int:
            out += b"\x02" + _PACK_Q(e & _MASK64)
        elif t is Fp2:
            out += b"\x03" + _PACK_Q(e.real & _MASK64) + _PACK_Q(e.imag & _MASK64)
        elif isinstance(e, str):
            b = e.encode("utf-8")
            out += b"\x04" + _PACK_I(len(b)) + b
        elif e is None:
            out += b"\x05"
        else:
            raise TypeError("unsupported FS element: %r" % (t,))
    return bytes(out)


def tree_tokens(obj) -> bytes:
    """Token form of a proof object for the library's proof writer
    (tp_pickle_tree): ints, strs, None, Fp2, lists and tuples; a list of
    non-negative u64 ints (two or more) becomes one packed int array (of
    bytes when every int is below 256)."""
    out = bytearray()

    def enc(x):
        t = type(x)
        if t is int:
            if not 0 <= x <= _MASK64:
                raise ValueError("proof ints are non-negative u64")
            out.extend(b"i" + _PACK_Q(x))
        elif x is None:
            out.extend(b"n")
        elif t is str:
            b = x.encode("utf-8")
            out.extend(b"s" + _PACK_I(len(b)) + b)
        elif t is Fp2:
            out.extend(b"f" + _PACK_Q(x.real) + _PACK_Q(x.imag))
        elif t is list or t is tuple:
            if (t is list and len(x) > 1 and all(type(v) is int and 0 <= v <= _MASK64 for v in x)):
                if max(x) < 256:
                    out.extend(b"b" + _PACK_I(len(x)) + bytes(x))
                else:
                    out.extend(b"a" + _PACK_I(len(x)) + struct.pack("<%dQ" % len(x), *x))
                return
            out.extend((b"l" if t is list else b"t") + _PACK_I(len(x)))
            for v in x:
                enc(v)
        else:
            raise TypeError("unsupported proof element %r" % (t,))
    enc(obj)
    return bytes(out)


def native_dumps(obj, lib) -> bytes:
    """pickle.dumps(obj, 5) through the library's proof writer."""
    import ctypes
    tok = tree_tokens(obj)
    ptr = ctypes.POINTER(ctypes.c_uint8)()
    n = ctypes.c_size_t()
    if lib.tp_pickle_tree(tok, len(tok), ctypes.byref(ptr), ctypes.byref(n)) != 0:
        raise ValueError("tp_pickle_tree rejected the tree")
    try:
        return ctypes.string_at(ptr, n.value)
    finally:
        lib.tp_free(ptr)


class _ProofUnpickler(pickle.Unpickler):
    """Admits only the Fp2 class; any other global in a proof is rejected."""

    def find_class(self, module, name):
        if (module, name) == ("libTruePix.extfield", "Fp2"):
            return Fp2
        raise pickle.UnpicklingError("proof references forbidden global %s.%s" % (module, name))


class FiatShamir:
    VERSION = 4
    ndb = None
    rvstart = None
    rvend = None
    wDiv = None

    def __init__(self, q: int = P, protocol: str = "truepix"):
        if q != P:
            raise ValueError("this transcript is specialised to q = 2^61 - 1")
        self.q = q
        self.protocol = protocol
        self.mask = P
        self.comp = q
        self.rnum = 0
        self.io = deque()
        self.trans = deque()
        self.can_put = True
        self.hash = hashlib.sha256()
        self.hash.update(b"TruePix," + str(protocol).encode("utf-8") + b",binhash1,")

    # absorption -------------------------------------------------------------
    def hash_update(self, vals) -> None:
        self.hash.update(encode_items(vals))

    def hash_append(self, elm) -> None:
        self.hash.update(encode_items([elm]))

    def absorb_int_array(self, arr) -> None:
        """Absorb a flat list of non-negative u64 ints held as a numpy array
        (same bytes as put(list_of_ints) without building the list)."""
        import numpy as np
        a = np.ascontiguousarray(arr, dtype="<u8")
        self.hash.update(b"\x01" + _PACK_I(a.size))
        self.hash.update(a.tobytes())

    # challenges ---------------------------------------------------------------
    def rand_scalar(self) -> int:
        ret = self.comp
        while ret >= self.comp:
            self.hash_append(self.rnum)
            self.rnum += 1
            ret = int.from_bytes(self.hash.digest()[-8:], "big") & self.mask
        return ret % self.q

    def rand_ext(self, label: str = "chal") -> Fp2:
        self.hash_append("ext")
        self.hash_append(label)
        self.hash_append("real")
        re = self.rand_scalar()
        self.hash_append("ext")
        self.hash_append(label)
        self.hash_append("imag")
        im = self.rand_scalar()
        return _raw(re, im)

    # transcript values --------------------------------------------------------
    def put(self, vals, is_io: bool = False) -> None:
        if not self.can_put:
            raise AssertionError("Cannot put values into an input FS transcript")
        is_dict = isinstance(vals, dict)
        if not is_dict and not isinstance(vals, list):
            vals = [vals]
        (self.io if is_io else self.trans).append(vals)
        self.hash_update(list(vals.items()) if is_dict else vals)

    def put_io_int_array(self, arr) -> None:
        """put(list(arr), True) for a large int array, hashing from numpy."""
        lst = arr.tolist()
        self.io.append(lst)
        self.absorb_int_array(arr)

    def take(self, take_io: bool = False):
        if self.can_put:
            raise AssertionError("Cannot take values from an output FS transcript")
        vals = (self.io if take_io else self.trans).popleft()
        self.hash_update(list(vals.items()) if isinstance(vals, dict) else vals)
        return vals

    # proof bytes --------------------------------------------------------------
    def to_string(self, full: bool = True) -> bytes:
        with proof_pickle_names():
            if full:
                return pickle.dumps((self.q, list(self.io), list(self.trans), self.ndb,
                                     self.rvstart, self.rvend, self.wDiv, self.VERSION), -1)
            return pickle.dumps(list(self.trans), -1)

    @classmethod
    def unpack_proof(cls, data: bytes):
        obj = _ProofUnpickler(_io.BytesIO(data)).load()
        (q, iovals, tvals, ndb, rvstart, rvend, wdiv, version) = obj
        if version != cls.VERSION:
            raise ValueError("Version: expected %d, got %d" % (cls.VERSION, version))
        return (q, iovals, tvals, ndb, rvstart, rvend, wdiv)

    @classmethod
    def from_string(cls, data: bytes, protocol: str = "truepix") -> "FiatShamir":
        (q, iovals, tvals, ndb, rvstart, rvend, wdiv) = cls.unpack_proof(data)
        fs = cls(q, protocol=protocol)
        fs.can_put = False
        fs.io.extend(iovals)
        fs.trans.extend(tvals)
        fs.ndb, fs.rvstart, fs.rvend, fs.wDiv = ndb, rvstart, rvend, wdiv
        return fs

    @classmethod
    def proof_size(cls, data: bytes):
        (_, _, tvals, _, _, _, _) = cls.unpack_proof(data)
        nelems = sum(1 for _ in flatten_iter(tvals))
        with proof_pickle_names():
            size = len(bz2.compress(pickle.dumps(tvals, -1)))
        return (nelems, size)
