"""ctypes loader for the CPU checker.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker or as the timed CPU baseline.  The product package
(truepix_amd) never imports it.

liboracle_fp2.so is the restated oracle (oracle/fp2_oracle.c).  The optional
oracle/_ref/*.so are the reference's own sources compiled in place by
`make -C oracle ref` (they exist only where /root/reference does).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
P = (1 << 61) - 1

_U64P = ctypes.POINTER(ctypes.c_uint64)
_I32P = ctypes.POINTER(ctypes.c_int)
_U64 = ctypes.c_uint64
_SZ = ctypes.c_size_t
_INT = ctypes.c_int


def _p(a: np.ndarray):
    """Pointer to a contiguous uint64 / int32 numpy array."""
    if a.dtype == np.int32:
        return a.ctypes.data_as(_I32P)
    assert a.dtype == np.uint64 and a.flags.c_contiguous
    return a.ctypes.data_as(_U64P)


def _bind(lib, name, argtypes, restype=None):
    fn = getattr(lib, name)
    fn.argtypes = argtypes
    fn.restype = restype
    return fn


class Oracle:
    """Restated oracle: oracle/fp2_oracle.c."""

    def __init__(self, path: str | None = None):
        path = path or os.path.join(HERE, "liboracle_fp2.so")
        if not os.path.isfile(path):
            raise FileNotFoundError("oracle library not built: %s (run make -C oracle)" % path)
        lib = ctypes.CDLL(path)
        self.lib = lib
        self.fp2_mul = _bind(lib, "or_fp2_mul", [_U64, _U64, _U64, _U64, _U64P])
        self.fp2_inv = _bind(lib, "or_fp2_inv", [_U64, _U64, _U64P])
        self.gatefn = _bind(lib, "or_gatefn", [_INT, _U64, _U64, _U64, _U64, _U64P])
        self.fold = _bind(lib, "or_fold", [_U64P, _U64P, _SZ, _U64, _U64, _U64P, _U64P])
        self.expand = _bind(lib, "or_expand", [_U64P, _U64P, _SZ, _SZ, _U64P, _U64P])
        self.sumcheck_early = _bind(lib, "or_sumcheck_early",
                                    [_INT, _I32P, _I32P, _I32P, _U64P, _U64P, _INT, _INT]
                                    + [_U64P] * 12 + [_U64P])
        self.compute_beta = _bind(lib, "or_compute_beta", [_U64P, _U64P, _INT, _U64, _U64, _U64P, _U64P])
        self.mle_eval_base = _bind(lib, "or_mle_eval_base", [_U64P, _U64P, _U64P, _INT, _U64P, _U64P])
        self.early_claim = _bind(lib, "or_early_claim",
                                 [_INT, _I32P, _I32P, _I32P, _U64P, _U64P, _INT,
                                  _U64P, _U64P, _U64P, _U64P, _U64P])
        self.early_fold = _bind(lib, "or_early_fold",
                                [_INT, _INT, _U64P, _U64P, _U64P, _U64P, _U64, _U64,
                                 _U64P, _U64P, _U64P, _U64P])
        self.sc2_round = _bind(lib, "or_sc2_round", [_U64P, _U64P, _SZ, _INT, _U64P])
        self.sc2_fold = _bind(lib, "or_sc2_fold", [_U64P, _U64P, _SZ, _U64, _U64, _INT])
        self.root_of_unity = _bind(lib, "or_root_of_unity", [_INT, _U64P])
        self.fft_raw = _bind(lib, "or_fft", [_U64P, _INT, _INT, _U64P])
        self.poly_eval_raw = _bind(lib, "or_poly_eval", [_U64P, _SZ, _U64, _U64, _U64P])
        self.pow_raw = _bind(lib, "or_fp2_pow", [_U64, _U64, _U64, _U64P])
        self.glibc_rand_raw = _bind(lib, "or_glibc_rand",
                                    [ctypes.c_uint32, _INT, ctypes.POINTER(ctypes.c_int32)])
        self.fft_gkr_raw = _bind(lib, "or_fft_gkr", [_INT, _U64P, _INT, _I32P, _U64P, _U64P, _I32P],
                                 _INT)

    # -- numpy conveniences -------------------------------------------------
    def mul(self, a, b):
        out = (ctypes.c_uint64 * 2)()
        self.fp2_mul(a[0], a[1], b[0], b[1], out)
        return (out[0], out[1])

    def root(self, log_order):
        out = (ctypes.c_uint64 * 2)()
        self.root_of_unity(log_order, out)
        return (out[0], out[1])

    def pow(self, x, e):
        out = (ctypes.c_uint64 * 2)()
        self.pow_raw(x[0], x[1], e, out)
        return (out[0], out[1])

    def poly_eval(self, coef: np.ndarray, x):
        coef = np.ascontiguousarray(coef, dtype=np.uint64)
        out = (ctypes.c_uint64 * 2)()
        self.poly_eval_raw(_p(coef), coef.size // 2, x[0], x[1], out)
        return (out[0], out[1])

    def fft(self, coef: np.ndarray, log_coef: int, log_order: int) -> np.ndarray:
        coef = np.ascontiguousarray(coef, dtype=np.uint64)
        assert coef.size == 2 << log_coef
        out = np.zeros(2 << log_order, dtype=np.uint64)
        self.fft_raw(_p(coef), log_coef, log_order, _p(out))
        return out

    def sc2_prove(self, f: np.ndarray, g: np.ndarray, challenge_fn, nthreads: int = 1):
        """Degree-2 sumcheck over interleaved tables; challenge_fn(round, poly6) -> (re, im).

        Returns (polys [nv x 6], challenges [nv x 2], final [4])."""
        f = np.array(f, dtype=np.uint64, copy=True)
        g = np.array(g, dtype=np.uint64, copy=True)
        n = f.size // 2
        nv = n.bit_length() - 1
        polys = np.zeros((nv, 6), dtype=np.uint64)
        chals = np.zeros((nv, 2), dtype=np.uint64)
        out6 = (ctypes.c_uint64 * 6)()
        for r in range(nv):
            self.sc2_round(_p(f), _p(g), n, nthreads, out6)
            polys[r] = list(out6)
            cr, ci = challenge_fn(r, tuple(out6))
            chals[r] = (cr, ci)
            self.sc2_fold(_p(f), _p(g), n, cr, ci, nthreads)
            n //= 2
        return polys, chals, np.array([f[0], f[1], g[0], g[1]], dtype=np.uint64)


def fft_gkr_result(fn, lg: int, with_polys: bool = True):
    """Runs an or_fft_gkr-shaped function; returns ok, polys (n x 6 limbs),
    claims (7 x 2), layers (2^lg last IFFT, 2^lg scaled, 64 sums; x 2 limbs)
    and proof_size."""
    n = 1 << lg
    n_rounds = (lg + 6) + lg + 2 * lg * lg
    polys = np.zeros(6 * n_rounds, np.uint64) if with_polys else None
    claims = np.zeros(14, np.uint64)
    layers = np.zeros(4 * n + 128, np.uint64)
    npol = ctypes.c_int(0)
    psize = ctypes.c_int(0)
    ok = fn(lg, _p(polys) if with_polys else None, n_rounds if with_polys else 0,
            ctypes.byref(npol), _p(claims), _p(layers), ctypes.byref(psize))
    return {"ok": ok, "polys": polys, "n_polys": npol.value, "claims": claims, "layers": layers,
            "proof_size": psize.value}


def _oracle_fft_gkr(self, lg: int, with_polys: bool = True):
    return fft_gkr_result(self.fft_gkr_raw, lg, with_polys)


def _oracle_glibc_rand(self, seed: int, n: int):
    out = (ctypes.c_int32 * n)()
    self.glibc_rand_raw(seed, n, out)
    return list(out)


Oracle.fft_gkr = _oracle_fft_gkr
Oracle.glibc_rand = _oracle_glibc_rand


def _orion_commit(fn, inputs, seed: int):
    arr = np.ascontiguousarray(inputs, dtype=np.uint8)
    a = ctypes.create_string_buffer(65)
    b = ctypes.create_string_buffer(65)
    cw = ctypes.c_longlong()
    rc = fn(int(arr.shape[0]), arr.tobytes(), int(seed) & (2 ** 64 - 1), a, b, ctypes.byref(cw))
    if rc != 0:
        raise ValueError("orion commit rejected copies=%d" % arr.shape[0])
    return a.value.decode(), b.value.decode(), cw.value


def _orion_encode(fn, col):
    """Codeword of one column [n][2] (re, im) -> [cw][2]."""
    src = np.ascontiguousarray(col, dtype=np.uint64)
    n = src.shape[0]
    dst = np.zeros((int(n * 1.72) + 16) * 2, np.uint64)
    cw = fn(n, _p(src), _p(dst))
    return dst[:2 * cw].reshape(cw, 2).copy()


_COMMIT_ARGS = [_INT, ctypes.c_char_p, _U64, ctypes.c_char_p, ctypes.c_char_p,
                ctypes.POINTER(ctypes.c_longlong)]


def _oracle_orion_set_layout(self, copy_size: int, const_idx: int, const_val: int):
    """The packing of a copy for every later commit and opening (default
    384, 384, 128)."""
    if _bind(self.lib, "or_orion_set_layout", [_INT, _INT, _INT], _INT)(copy_size, const_idx,
                                                                      const_val) != 0:
        raise ValueError("not a packing: %d %d %d" % (copy_size, const_idx, const_val))


def _oracle_orion_commit(self, inputs, seed: int):
    """(masked_root, mask_root, cw_len) of the signer's commitment."""
    if not hasattr(self, "_orc"):
        self._orc = _bind(self.lib, "or_orion_commit", _COMMIT_ARGS, _INT)
    return _orion_commit(self._orc, inputs, seed)


def _oracle_orion_encode(self, col):
    if not hasattr(self, "_ore"):
        self._ore = _bind(self.lib, "or_orion_encode", [ctypes.c_longlong, _U64P, _U64P],
                          ctypes.c_longlong)
    return _orion_encode(self._ore, col)


def _oracle_sha256(self, data: bytes) -> bytes:
    if not hasattr(self, "_sha"):
        self._sha = _bind(self.lib, "or_sha256", [ctypes.c_char_p, _SZ, ctypes.c_char_p])
    out = ctypes.create_string_buffer(32)
    self._sha(data, len(data), out)
    return out.raw


_OPEN_ARGS = [_INT, ctypes.c_char_p, _U64, _INT, _U64P, _INT, ctypes.POINTER(ctypes.c_longlong), _INT,
              ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(_SZ)]


def _orion_open(fn, free, inputs, seed, use_group1, r, col_q, with_answer):
    """OpeningProof bytes (and the answer when the function returns it)."""
    arr = np.ascontiguousarray(inputs, dtype=np.uint8)
    rr = np.ascontiguousarray(np.asarray(r, dtype=np.uint64).reshape(-1))
    q = np.ascontiguousarray(col_q, dtype=np.int64)
    buf = ctypes.c_void_p()
    n = _SZ()
    extra = []
    ans = np.zeros(2, np.uint64)
    if with_answer:
        extra = [_p(ans)]
    rc = fn(int(arr.shape[0]), arr.tobytes(), int(seed) & (2 ** 64 - 1), 1 if use_group1 else 0,
            _p(rr), rr.size // 2, q.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), q.size,
            ctypes.byref(buf), ctypes.byref(n), *extra)
    if rc != 0:
        raise ValueError("orion open rejected")
    data = ctypes.string_at(buf.value, n.value)
    free(buf)
    return (data, (int(ans[0]), int(ans[1]))) if with_answer else data


def _oracle_orion_open(self, inputs, seed: int, r, col_q, use_group1: bool = True):
    """(OpeningProof bytes, answer) of one committed polynomial."""
    if not hasattr(self, "_oro"):
        self._oro = _bind(self.lib, "or_orion_open", _OPEN_ARGS + [_U64P], _INT)
        self._orf = _bind(self.lib, "or_free", [ctypes.c_void_p])
    return _orion_open(self._oro, self._orf, inputs, seed, use_group1, r, col_q, True)


_DRAW_ARGS = [ctypes.c_char_p, ctypes.c_longlong, _INT, ctypes.c_longlong, _U64P, _INT,
              ctypes.POINTER(ctypes.c_longlong)]


def _draw(fn, root: bytes, n: int, use_group1: bool, cw_len: int, r):
    rr = np.ascontiguousarray(np.asarray(r, dtype=np.uint64).reshape(-1))
    q = np.zeros(4096, np.int64)
    k = fn(root, n, 1 if use_group1 else 0, cw_len, _p(rr), rr.size // 2,
           q.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
    return [int(x) for x in q[:k]]


def _oracle_orion_draw_queries(self, root: bytes, n: int, use_group1: bool, cw_len: int, r):
    """The code positions the opening's NIZK transcript draws (the opened
    root as digest bytes, N, the group, the point r)."""
    if not hasattr(self, "_ord"):
        self._ord = _bind(self.lib, "or_orion_draw_queries", _DRAW_ARGS, _INT)
    return _draw(self._ord, root, n, use_group1, cw_len, r)


Oracle.orion_open = _oracle_orion_open
Oracle.orion_draw_queries = _oracle_orion_draw_queries
Oracle.orion_commit = _oracle_orion_commit
Oracle.orion_set_layout = _oracle_orion_set_layout
Oracle.orion_encode = _oracle_orion_encode
Oracle.sha256 = _oracle_sha256


_VIRGO_ARGS = [ctypes.c_char_p, ctypes.c_longlong, _INT, _U64P, _INT, ctypes.c_char_p, _U64P,
               ctypes.c_longlong, _U64P, ctypes.c_longlong]


def _virgo(fn, com_root: bytes, n: int, use_group1: bool, r, combined_root: bytes, message):
    """Words of the Virgo code-switching transcript (layout in
    oracle/ref_harness_virgo.cpp)."""
    rr = np.ascontiguousarray(np.asarray(r, dtype=np.uint64).reshape(-1))
    msg = np.ascontiguousarray(np.asarray(message, dtype=np.uint64).reshape(-1))
    cap = 1 << 16
    while True:
        out = np.zeros(cap, np.uint64)
        k = fn(com_root, n, 1 if use_group1 else 0, _p(rr), rr.size // 2, combined_root, _p(msg),
               msg.size // 2, _p(out), cap)
        if k < 0:
            raise ValueError("Virgo transcript refused its arguments")
        if k <= cap:
            return out[:k]
        cap = k


def _oracle_orion_virgo(self, com_root: bytes, n: int, use_group1: bool, r, combined_root: bytes,
                        message):
    if not hasattr(self, "_virgo_fn"):
        self._virgo_fn = _bind(self.lib, "or_orion_virgo", _VIRGO_ARGS, ctypes.c_longlong)
    return _virgo(self._virgo_fn, com_root, n, use_group1, r, combined_root, message)


def _oracle_orion_virgo_plant(self, k: int):
    """Plants a mistake in the next transcripts (0: none), for the tests of
    the verifier's checks."""
    _bind(self.lib, "or_orion_virgo_plant", [_INT])(k)


Oracle.orion_virgo = _oracle_orion_virgo
Oracle.orion_virgo_plant = _oracle_orion_virgo_plant


class RefVirgo:
    """The Virgo code-switching proof of the reference's NIZK opening, run by
    the reference's own generate_circuit, zk_prover and zk_verifier pieces
    (oracle/ref_harness_virgo.cpp)."""

    PATH = os.path.join(HERE, "_ref", "libvirgo_ref.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.isfile(cls.PATH)

    def __init__(self):
        self.fn = _bind(ctypes.CDLL(self.PATH), "ref_orion_virgo", _VIRGO_ARGS, ctypes.c_longlong)

    def transcript(self, com_root: bytes, n: int, use_group1: bool, r, combined_root: bytes, message):
        return _virgo(self.fn, com_root, n, use_group1, r, combined_root, message)


class RefOrion:
    """The signer's commitment and an opening computed with the reference's
    own layout, mask, expander code, row hash, Merkle tree, siblings and
    proof serializer (oracle/ref_harness_orion.cpp)."""

    PATH = os.path.join(HERE, "_ref", "liborion_ref.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.isfile(cls.PATH)

    def __init__(self):
        lib = ctypes.CDLL(self.PATH)
        self.c = _bind(lib, "ref_orion_commit", _COMMIT_ARGS, _INT)
        self.e = _bind(lib, "ref_orion_encode", [ctypes.c_longlong, _U64P, _U64P], ctypes.c_longlong)

    @staticmethod
    def set_layout(copy_size: int, const_idx: int, const_val: int):
        """The packing, through the environment variables the reference's
        make_layout reads (truepix_gkr.h:107-150) at every call."""
        os.environ["TRUEPIX_COPY_SIZE"] = str(copy_size)
        os.environ["TRUEPIX_CONST_IDX"] = str(const_idx)
        os.environ["TRUEPIX_CONST_VAL"] = str(const_val)

    def commit(self, inputs, seed: int):
        return _orion_commit(self.c, inputs, seed)

    def open(self, inputs, seed: int, r, col_q, use_group1: bool = True) -> bytes:
        if not hasattr(self, "o"):
            lib = ctypes.CDLL(self.PATH)
            self.o = _bind(lib, "ref_orion_open", _OPEN_ARGS, _INT)
            self.f = _bind(lib, "ref_free", [ctypes.c_void_p])
        return _orion_open(self.o, self.f, inputs, seed, use_group1, r, col_q, False)

    def draw_queries(self, root: bytes, n: int, use_group1: bool, cw_len: int, r):
        if not hasattr(self, "d"):
            self.d = _bind(ctypes.CDLL(self.PATH), "ref_orion_draw_queries", _DRAW_ARGS, _INT)
        return _draw(self.d, root, n, use_group1, cw_len, r)

    def encode(self, col):
        return _orion_encode(self.e, col)


class RefFftGkr:
    """fft_circuit_GKR.cpp compiled in place, driven through its own layer
    functions (oracle/ref_harness_fftgkr.cpp)."""

    PATH = os.path.join(HERE, "_ref", "libfftgkr_ref.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.isfile(cls.PATH)

    def __init__(self):
        lib = ctypes.CDLL(self.PATH)
        self.raw = _bind(lib, "ref_fft_gkr", [_INT, _U64P, _U64P, _I32P], _INT)

    def run(self, lg: int):
        n = 1 << lg
        claims = np.zeros(14, np.uint64)
        layers = np.zeros(4 * n + 128, np.uint64)
        psize = ctypes.c_int(0)
        ok = self.raw(lg, _p(claims), _p(layers), ctypes.byref(psize))
        return {"ok": ok, "claims": claims, "layers": layers, "proof_size": psize.value}


class RefNative:
    """TP/native_fp2/native_fp2.cpp compiled from source into oracle/_ref."""

    PATH = os.path.join(HERE, "_ref", "libnative_fp2_ref.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.isfile(cls.PATH)

    def __init__(self):
        lib = ctypes.CDLL(self.PATH)
        self.lib = lib
        self.fold = _bind(lib, "native_fp2_fold", [_U64P, _U64P, _SZ, _U64, _U64, _U64P, _U64P])
        self.expand = _bind(lib, "native_fp2_expand", [_U64P, _U64P, _SZ, _SZ, _U64P, _U64P])
        self.sumcheck_early = _bind(lib, "native_fp2_sumcheck_early",
                                    [_INT, _I32P, _I32P, _I32P, _U64P, _U64P, _INT, _INT]
                                    + [_U64P] * 12 + [_U64P])
        self.compute_beta = _bind(lib, "native_fp2_compute_beta", [_U64P, _U64P, _INT, _U64, _U64, _U64P, _U64P])
        self.mle_eval_base = _bind(lib, "native_fp2_mle_eval_base", [_U64P, _U64P, _U64P, _INT, _U64P, _U64P])
        self.early_create = _bind(lib, "native_fp2_early_create",
                                  [_INT, _I32P, _I32P, _I32P, _U64P, _U64P, _INT, _INT,
                                   _U64P, _U64P, _U64P, _U64P], ctypes.c_void_p)
        self.early_claim = _bind(lib, "native_fp2_early_claim", [ctypes.c_void_p, _U64P], _INT)
        self.early_fold = _bind(lib, "native_fp2_early_fold", [ctypes.c_void_p, _U64, _U64], _INT)
        self.early_finish = _bind(lib, "native_fp2_early_finish",
                                  [ctypes.c_void_p, _U64P, _U64P, _U64P, _U64P], _INT)
        self.early_destroy = _bind(lib, "native_fp2_early_destroy", [ctypes.c_void_p])


class RefSumcheck:
    """sumcheck_phase1_update from OZ/src/VPD/fft_circuit_GKR.cpp, compiled in place."""

    PATH = os.path.join(HERE, "_ref", "libsumcheck_ref.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.isfile(cls.PATH)

    def __init__(self):
        lib = ctypes.CDLL(self.PATH)
        self.sc2 = _bind(lib, "ref_sc2", [_U64P, _U64P, _INT, _U64P, _U64P, _U64P])

    def prove(self, f, g, chals):
        f = np.ascontiguousarray(f, dtype=np.uint64)
        g = np.ascontiguousarray(g, dtype=np.uint64)
        nv = (f.size // 2).bit_length() - 1
        ch = np.ascontiguousarray(chals, dtype=np.uint64).reshape(-1)
        polys = np.zeros(nv * 6, dtype=np.uint64)
        fin = np.zeros(4, dtype=np.uint64)
        self.sc2(_p(f), _p(g), nv, _p(ch), _p(polys), _p(fin))
        return polys.reshape(nv, 6), fin


class RefRS:
    """fast_fourier_transform / get_root_of_unity from Orion, compiled in place."""

    PATH = os.path.join(HERE, "_ref", "librs_ref.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.isfile(cls.PATH)

    def __init__(self):
        lib = ctypes.CDLL(self.PATH)
        self.root_raw = _bind(lib, "ref_root_of_unity", [_INT, _U64P])
        self.fft_raw = _bind(lib, "ref_fft", [_U64P, _INT, _INT, _U64P])

    def root(self, log_order):
        out = (ctypes.c_uint64 * 2)()
        self.root_raw(log_order, out)
        return (out[0], out[1])

    def fft(self, coef, log_coef, log_order):
        coef = np.ascontiguousarray(coef, dtype=np.uint64)
        out = np.zeros(2 << log_order, dtype=np.uint64)
        self.fft_raw(_p(coef), log_coef, log_order, _p(out))
        return out


def rand_fp2(rng: np.random.Generator, n: int) -> np.ndarray:
    """n uniform canonical Fp2 elements, interleaved (re, im), by rejection."""
    out = np.empty(2 * n, dtype=np.uint64)
    filled = 0
    while filled < 2 * n:
        x = rng.integers(0, 1 << 61, size=2 * n - filled + 64, dtype=np.uint64)
        x = x[x < P]
        take = min(x.size, 2 * n - filled)
        out[filled:filled + take] = x[:take]
        filled += take
    return out
