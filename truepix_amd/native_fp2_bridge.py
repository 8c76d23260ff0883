"""ctypes bridge to libtruepix_hip.so: the Python face of the HIP kernels.

Same function names, arguments and return shapes as the reference bridge
(TP/libTruePix/native_fp2_bridge.py:273-542): ``fold_values``,
``sumcheck_early``, ``EarlySession`` (claim / fold / finish / destroy),
``early_session_create``, ``resolve_gate_type_idx``, ``cache_gate_tables``,
``compute_beta``, ``mle_eval_base``, ``Fp2Buf``, ``AVAILABLE``,
``GPU_ACTIVE``, ``cuda_available()``, ``early_session_supported()``.

Differences, on purpose:
  * one library, libtruepix_hip.so, and no CUDA runtime preload;
  * no host fallback: if the library or the GPU is missing, AVAILABLE and
    GPU_ACTIVE are False and every call raises RuntimeError naming the cause
    (GPU_FAIL_REASON); a device error inside a call raises RuntimeError with
    the library's message;
  * Fp2Buf is backed by numpy uint64 arrays; packing uses numpy, not
    per-element Python loops.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

from truepix_amd.extfield import BASE_PRIME_61, Fp2, _raw

P = BASE_PRIME_61
HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libtruepix_hip.so")

_U64 = ctypes.c_uint64
_I32 = ctypes.c_int
_SIZE = ctypes.c_size_t
_VP = ctypes.c_void_p
_U64P = ctypes.POINTER(_U64)
_I32P = ctypes.POINTER(_I32)

ENABLE_GPU = os.environ.get("TRUEPIX_ENABLE_GPU", "").strip().lower() not in ("0", "false", "no", "off")

LIB = None          # the loaded CDLL (symbols usable even without a device)
AVAILABLE = False   # library loaded and a HIP device is active
GPU_ACTIVE = False
GPU_FAIL_REASON = ""


def _sig(lib, name, argtypes, restype=None):
    fn = getattr(lib, name)
    fn.argtypes = argtypes
    fn.restype = restype


class FrameJob(ctypes.Structure):
    """tp_frame_job (include/truepix_hip.h)."""
    _fields_ = [
        ("circuit", _VP), ("n_copies", _I32), ("n_private", _I32), ("input_bytes", _I32),
        ("inputs", _VP), ("pid", _U64), ("n_vole", _I32),
        ("vole_u_re", _U64P), ("vole_u_im", _U64P), ("vole_v_re", _U64P), ("vole_v_im", _U64P),
        ("n_binding", _I32), ("binding_kind", _I32P), ("binding_int", _U64P),
        ("binding_str", ctypes.POINTER(ctypes.c_char_p)), ("keep_outputs", _I32),
        ("inputs_on_device", _I32),
    ]


def bind(lib) -> None:
    """Declare every exported symbol's signature (include/truepix_hip.h)."""
    _sig(lib, "native_fp2_init", [], _I32)
    _sig(lib, "native_fp2_cuda_available", [], _I32)
    _sig(lib, "native_fp2_fold", [_U64P, _U64P, _SIZE, _U64, _U64, _U64P, _U64P])
    _sig(lib, "native_fp2_expand", [_U64P, _U64P, _SIZE, _SIZE, _U64P, _U64P])
    _sig(lib, "native_fp2_sumcheck_early",
         [_I32, _I32P, _I32P, _I32P, _U64P, _U64P, _I32, _I32] + [_U64P] * 12 + [_U64P])
    _sig(lib, "native_fp2_compute_beta", [_U64P, _U64P, _I32, _U64, _U64, _U64P, _U64P])
    _sig(lib, "native_fp2_mle_eval_base", [_U64P, _U64P, _U64P, _I32, _U64P, _U64P])
    _sig(lib, "native_fp2_early_create",
         [_I32, _I32P, _I32P, _I32P, _U64P, _U64P, _I32, _I32, _U64P, _U64P, _U64P, _U64P], _VP)
    _sig(lib, "native_fp2_early_claim", [_VP, _U64P], _I32)
    _sig(lib, "native_fp2_early_fold", [_VP, _U64, _U64], _I32)
    _sig(lib, "native_fp2_early_finish", [_VP, _U64P, _U64P, _U64P, _U64P], _I32)
    _sig(lib, "native_fp2_early_n_copies", [_VP], _I32)
    _sig(lib, "native_fp2_early_n_wires", [_VP], _I32)
    _sig(lib, "native_fp2_early_destroy", [_VP])
    _sig(lib, "tp_last_error", [], ctypes.c_char_p)
    _sig(lib, "tp_clear_error", [])
    _sig(lib, "tp_device_count", [], _I32)
    _sig(lib, "tp_sc2_prove", [_VP, _VP, _I32, _U64, _U64P, _U64P, _U64P, _U64P, _VP], _I32)
    _sig(lib, "tp_sc2_set_timing", [_I32])
    _sig(lib, "tp_sc2_set_ahead", [_I32])
    _sig(lib, "tp_sc2_set_block_pass", [_I32])
    _sig(lib, "tp_rs_encode", [_VP, _I32, _I32, _I32, _VP, _I32, _VP], _I32)
    _sig(lib, "tp_early_create_dev",
         [_I32, _I32P, _I32P, _I32P, _U64P, _U64P, _I32, _I32, _VP, _VP, _U64P, _U64P], _VP)
    _sig(lib, "tp_eval_layer", [_VP, _I32, _VP, _I32, _I32, _I32, _I32P, _I32P, _I32P, _I32P, _I32,
                                _VP], _I32)
    _sig(lib, "tp_wire_table", [_VP, _I32, _I32, _I32, _I32, _VP, _VP, _VP], _I32)
    _sig(lib, "tp_late_create", [_I32, _I32P, _I32P, _I32P, _U64P, _U64P, _I32, _U64P, _U64P,
                                 _U64, _U64], _VP)
    _sig(lib, "tp_late_round", [_VP, _I32, _U64P, _U64P], _I32)
    _sig(lib, "tp_late_values", [_VP, _U64P, _I32, _U64P], _I32)
    _sig(lib, "tp_late_destroy", [_VP])
    _sig(lib, "tp_wiring_mlext", [_I32, _I32P, _I32P, _I32P, _U64P, _U64P, _I32, _U64P, _U64P, _I32,
                                  _U64P, _U64P, _I32, _U64P], _I32)
    _sig(lib, "tp_mle_eval_base_dev", [_VP, _I32, _U64P, _U64P, _VP], _I32)
    _sig(lib, "tp_mle_eval_base_dev_noalloc", [_VP, _I32, _U64P, _U64P, _VP], _I32)
    _sig(lib, "tp_sc2_stats", [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                               _I32P])
    _sig(lib, "tp_sc2_tail_stats", [ctypes.POINTER(ctypes.c_double), _I32P])
    _sig(lib, "tp_sc2_class_stats", [ctypes.POINTER(ctypes.c_double), _I32])
    _sig(lib, "tp_sc2_tail_profile", [ctypes.POINTER(ctypes.c_double), _I32])
    _sig(lib, "tp_sc2_mid_profile", [ctypes.POINTER(ctypes.c_double), _I32])
    _sig(lib, "tp_sc2_schedule", [_I32P, _I32P, _I32P])
    _sig(lib, "tp_sc2_set_gate_ticks", [ctypes.c_longlong])
    _sig(lib, "tp_sc2_set_spin_ticks", [ctypes.c_longlong])
    _sig(lib, "tp_sc2_mid_fallbacks", [], _I32)
    _U8PP = ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8))
    _sig(lib, "tp_pickle_tree", [ctypes.c_char_p, _SIZE, _U8PP, ctypes.POINTER(_SIZE)], _I32)
    _sig(lib, "tp_free", [_VP])
    _sig(lib, "tp_sha256", [ctypes.c_char_p, _SIZE, _SIZE, ctypes.c_char_p])
    _sig(lib, "tp_sha256_has_ni", [], _I32)
    _D = ctypes.c_double
    _sig(lib, "tp_circuit_create", [_I32, _I32P, _I32P, _I32P, _I32P, _I32P, _I32, _I32P, _U64P],
         _VP)
    _sig(lib, "tp_circuit_destroy", [_VP])
    _sig(lib, "tp_orion_commit", [_I32, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_char_p,
                                  ctypes.c_char_p, ctypes.POINTER(_D)], _I32)
    _sig(lib, "tp_orion_encode", [ctypes.c_longlong, _I32, _U64P, _U64P], ctypes.c_longlong)
    _sig(lib, "tp_orion_set_layout", [_I32, _I32, _I32], _I32)
    _sig(lib, "tp_orion_open", [_I32, ctypes.c_char_p, ctypes.c_uint64, _I32, _U64P, _I32, _VP, _I32,
                                ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t),
                                _U64P], _I32)
    _sig(lib, "tp_orion_open_nizk", [_I32, ctypes.c_char_p, ctypes.c_uint64, _I32, _U64P, _I32,
                                     ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t),
                                     _U64P], _I32)
    _LL = ctypes.c_longlong
    _sig(lib, "tp_orion_open_nizk_virgo", [_I32, ctypes.c_char_p, ctypes.c_uint64, _I32, _U64P, _I32,
                                           ctypes.POINTER(ctypes.c_void_p),
                                           ctypes.POINTER(ctypes.c_size_t), _U64P, _U64P, _LL,
                                           ctypes.POINTER(_LL)], _I32)
    _sig(lib, "tp_orion_virgo_stats", [ctypes.POINTER(_D), _I32P, _I32P], _I32)
    _sig(lib, "tp_yuv_pack", [_VP, _VP, _VP, _I32, _I32, _LL, _LL, _VP, _I32, _VP, _VP], _LL)
    _sig(lib, "tp_yuv_unpack", [_VP, _LL, _I32, _I32, _LL, _LL, _VP, _VP, _VP, _VP], _I32)
    _sig(lib, "tp_yuv_pack_host", [_VP, _VP, _VP, _I32, _I32, _VP, _I32, _VP], _LL)
    _sig(lib, "tp_yuv_unpack_host", [_VP, _LL, _I32, _I32, _VP, _VP, _VP], _I32)
    _sig(lib, "tp_pool_create", [_I32], _VP)
    _sig(lib, "tp_pool_create_batched", [_I32, _I32, ctypes.c_uint64, _I32], _VP)
    _sig(lib, "tp_pool_job_stats", [_VP, ctypes.c_int64, _I32P, _U64P], _I32)
    _sig(lib, "tp_pool_proof_view", [_VP, ctypes.c_int64, ctypes.POINTER(_VP),
                                     ctypes.POINTER(_SIZE)], _I32)
    _sig(lib, "tp_pool_memory", [_VP, _U64P, _U64P, _U64P, _U64P, _U64P], _I32)
    _sig(lib, "tp_pool_batch_profile", [_VP, ctypes.POINTER(_D), _I32], _I32)
    _sig(lib, "tp_pool_tuning", [_VP, ctypes.POINTER(_D), _I32], _I32)
    _sig(lib, "tp_pool_fix_schedule", [_VP, ctypes.c_longlong, _I32], _I32)
    _sig(lib, "tp_pool_destroy", [_VP])
    _sig(lib, "tp_pool_now", [_VP], _D)
    _sig(lib, "tp_pool_submit", [_VP, ctypes.POINTER(FrameJob)], ctypes.c_int64)
    _sig(lib, "tp_pool_submit_many", [_VP, ctypes.POINTER(FrameJob), _I32,
                                      ctypes.POINTER(ctypes.c_int64)], _I32)
    _sig(lib, "tp_pool_wait", [_VP, ctypes.c_int64, _D], _I32)
    _sig(lib, "tp_pool_error", [_VP, ctypes.c_int64], ctypes.c_char_p)
    _sig(lib, "tp_pool_proof", [_VP, ctypes.c_int64, _VP, _SIZE, ctypes.POINTER(_SIZE)], _I32)
    _sig(lib, "tp_pool_info", [_VP, ctypes.c_int64, _U64P, _I32, _I32P, _U64P,
                               ctypes.POINTER(_D), ctypes.POINTER(_D)], _I32)
    _sig(lib, "tp_pool_outputs", [_VP, ctypes.c_int64, _U64P, _SIZE, ctypes.POINTER(_SIZE)], _I32)
    _sig(lib, "tp_pool_release", [_VP, ctypes.c_int64])
    _sig(lib, "tp_fft_gkr", [_I32, ctypes.POINTER(_D), _I32P, ctypes.POINTER(_D)], _I32)
    _sig(lib, "tp_fft_gkr_trace", [_I32, _U64P, _I32, _I32P, _U64P, _U64P, _I32P], _I32)
    _sig(lib, "tp_fft_gkr_set_graph", [_I32])


def load_library(path: str = LIB_PATH):
    """dlopen the library and bind signatures (works without a GPU)."""
    if not os.path.isfile(path):
        raise FileNotFoundError("libtruepix_hip.so is not built (%s); run "
                                "python -m truepix_amd.build" % path)
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    bind(lib)
    return lib


def _load():
    global LIB, AVAILABLE, GPU_ACTIVE, GPU_FAIL_REASON
    try:
        LIB = load_library()
    except (OSError, FileNotFoundError, AttributeError) as exc:
        GPU_FAIL_REASON = str(exc)
        return
    if not ENABLE_GPU:
        GPU_FAIL_REASON = "TRUEPIX_ENABLE_GPU disabled (there is no host fallback)"
        return
    if LIB.tp_device_count() <= 0:
        GPU_FAIL_REASON = "no HIP device visible"
        return
    if LIB.native_fp2_init() != 0 or LIB.native_fp2_cuda_available() != 1:
        GPU_FAIL_REASON = "device init failed: %s" % last_error()
        return
    AVAILABLE = True
    GPU_ACTIVE = True


def last_error() -> str:
    if LIB is None:
        return GPU_FAIL_REASON
    msg = LIB.tp_last_error()
    return msg.decode() if msg else ""


def require() -> None:
    if not AVAILABLE:
        raise RuntimeError("truepix HIP library unavailable: %s" % GPU_FAIL_REASON)


def check(what: str) -> None:
    """Raise if the last library call recorded an error."""
    msg = last_error()
    if msg:
        raise RuntimeError("%s failed: %s" % (what, msg))


_load()


def cuda_available() -> bool:
    return GPU_ACTIVE


def early_session_supported() -> bool:
    return GPU_ACTIVE


# --- buffers ---------------------------------------------------------------

def _ptr(a: np.ndarray):
    if a.dtype == np.int32:
        return a.ctypes.data_as(_I32P)
    return a.ctypes.data_as(_U64P)


class Fp2Buf:
    """Contiguous Fp2 vector (real[], imag[]) with list-like access."""
    __slots__ = ("real", "imag", "n")

    def __init__(self, real, imag, n: Optional[int] = None):
        self.real = np.ascontiguousarray(real, dtype=np.uint64)
        self.imag = np.ascontiguousarray(imag, dtype=np.uint64)
        self.n = int(n if n is not None else self.real.size)

    def __len__(self) -> int:
        return self.n

    def __getitem__(self, idx):
        if isinstance(idx, slice):
            start, stop, step = idx.indices(self.n)
            if step != 1:
                return [_raw(int(self.real[i]), int(self.imag[i])) for i in range(start, stop, step)]
            return Fp2Buf(self.real[start:stop].copy(), self.imag[start:stop].copy(),
                          max(0, stop - start))
        if idx < 0:
            idx += self.n
        return _raw(int(self.real[idx]), int(self.imag[idx]))

    def __iter__(self):
        re, im = self.real.tolist(), self.imag.tolist()
        for i in range(self.n):
            yield _raw(re[i], im[i])

    def to_list(self) -> List[Fp2]:
        return list(self)


def as_coords(values) -> Tuple[np.ndarray, np.ndarray]:
    """(re, im) uint64 arrays from an Fp2Buf, a sequence of Fp2/int, or arrays."""
    if isinstance(values, Fp2Buf):
        return values.real[:values.n], values.imag[:values.n]
    n = len(values)
    re = np.zeros(n, dtype=np.uint64)
    im = np.zeros(n, dtype=np.uint64)
    if n and type(values[0]) is int and all(type(v) is int for v in values):
        re[:] = np.array([v % P for v in values], dtype=np.uint64)
        return re, im
    for i, v in enumerate(values):
        if type(v) is Fp2:
            re[i] = v.real
            im[i] = v.imag
        else:
            re[i] = int(v) % P
    return re, im


def _pack_rows(rows, n_wires: int, n_cols: int):
    re = np.zeros((n_wires, n_cols), dtype=np.uint64)
    im = np.zeros((n_wires, n_cols), dtype=np.uint64)
    for w in range(n_wires):
        r, i = as_coords(rows[w])
        re[w, :n_cols] = r[:n_cols]
        im[w, :n_cols] = i[:n_cols]
    return re, im


def _chal(val) -> Fp2:
    if type(val) is Fp2:
        return val
    return _raw(int(val) % P, 0)


def _gate_arrays(gate_type, in0, in1, z1):
    gt = np.ascontiguousarray(gate_type, dtype=np.int32)
    i0 = np.ascontiguousarray(in0, dtype=np.int32)
    i1 = np.ascontiguousarray(in1, dtype=np.int32)
    zr, zi = as_coords(list(z1) if not isinstance(z1, Fp2Buf) else z1)
    return gt, i0, i1, np.ascontiguousarray(zr), np.ascontiguousarray(zi)


def _out4(out8) -> List[Fp2]:
    return [_raw(int(out8[2 * t]), int(out8[2 * t + 1])) for t in range(4)]


# --- batch functions -----------------------------------------------------------

def fold_values(values, challenge, expand_ncopies: int = 1) -> Tuple[Fp2Buf, Fp2Buf]:
    """Fold pairs; returns (outputs, scratch) (TP/libTruePix/native_fp2_bridge.py:273-295)."""
    require()
    chal = _chal(challenge)
    in_r, in_i = as_coords(values)
    in_r, in_i = np.ascontiguousarray(in_r), np.ascontiguousarray(in_i)
    n_out = in_r.size // 2
    out_r = np.empty(n_out, dtype=np.uint64)
    out_i = np.empty(n_out, dtype=np.uint64)
    LIB.native_fp2_fold(_ptr(in_r), _ptr(in_i), n_out, chal.real, chal.imag, _ptr(out_r), _ptr(out_i))
    check("native_fp2_fold")
    scratch = Fp2Buf(out_r, out_i, n_out)
    if expand_ncopies <= 1:
        return scratch, scratch
    n_exp = n_out * expand_ncopies
    exp_r = np.empty(n_exp, dtype=np.uint64)
    exp_i = np.empty(n_exp, dtype=np.uint64)
    LIB.native_fp2_expand(_ptr(out_r), _ptr(out_i), n_out, expand_ncopies, _ptr(exp_r), _ptr(exp_i))
    check("native_fp2_expand")
    return Fp2Buf(exp_r, exp_i, n_exp), scratch


def sumcheck_early(gate_type, in0, in1, z1, n_copies: int, n_wires: int, values, fact0, fact1,
                   beta, beta_f0, beta_f1) -> List[Fp2]:
    """Early-round claim with explicit -1/2 tables (native_fp2_bridge.py:298-350)."""
    require()
    gt, i0, i1, zr, zi = _gate_arrays(gate_type, in0, in1, z1)
    half = n_copies >> 1
    vr, vi = _pack_rows(values, n_wires, n_copies)
    f0r, f0i = _pack_rows(fact0, n_wires, half)
    f1r, f1i = _pack_rows(fact1, n_wires, half)
    br, bi = (np.ascontiguousarray(x) for x in as_coords(beta))
    b0r, b0i = (np.ascontiguousarray(x) for x in as_coords(beta_f0))
    b1r, b1i = (np.ascontiguousarray(x) for x in as_coords(beta_f1))
    out8 = np.zeros(8, dtype=np.uint64)
    LIB.native_fp2_sumcheck_early(
        len(gt), _ptr(gt), _ptr(i0), _ptr(i1), _ptr(zr), _ptr(zi), n_copies, n_wires,
        _ptr(vr), _ptr(vi), _ptr(f0r), _ptr(f0i), _ptr(f1r), _ptr(f1i),
        _ptr(br), _ptr(bi), _ptr(b0r), _ptr(b0i), _ptr(b1r), _ptr(b1i), _ptr(out8))
    check("native_fp2_sumcheck_early")
    return _out4(out8)


class EarlySession:
    """Resident early-round GKR session (native_fp2_bridge.py:353-410)."""

    __slots__ = ("_handle", "_n_wires")

    def __init__(self, handle, n_wires: int):
        self._handle = handle
        self._n_wires = int(n_wires)

    @property
    def n_copies(self) -> int:
        return int(LIB.native_fp2_early_n_copies(self._handle))

    @property
    def n_wires(self) -> int:
        return self._n_wires

    def claim(self) -> List[Fp2]:
        out8 = np.zeros(8, dtype=np.uint64)
        if LIB.native_fp2_early_claim(self._handle, _ptr(out8)) != 0:
            raise RuntimeError("native_fp2_early_claim failed: %s" % last_error())
        return _out4(out8)

    def fold(self, challenge) -> None:
        chal = _chal(challenge)
        if LIB.native_fp2_early_fold(self._handle, chal.real, chal.imag) != 0:
            raise RuntimeError("native_fp2_early_fold failed: %s" % last_error())

    def finish_arrays(self) -> Tuple[np.ndarray, np.ndarray, Fp2]:
        n = self._n_wires
        wr = np.zeros(n, dtype=np.uint64)
        wi = np.zeros(n, dtype=np.uint64)
        br = np.zeros(1, dtype=np.uint64)
        bi = np.zeros(1, dtype=np.uint64)
        if LIB.native_fp2_early_finish(self._handle, _ptr(wr), _ptr(wi), _ptr(br), _ptr(bi)) != 0:
            raise RuntimeError("native_fp2_early_finish failed: %s" % last_error())
        return wr, wi, _raw(int(br[0]), int(bi[0]))

    def finish(self) -> Tuple[List[Fp2], Fp2]:
        wr, wi, beta = self.finish_arrays()
        re, im = wr.tolist(), wi.tolist()
        return [_raw(re[i], im[i]) for i in range(self._n_wires)], beta

    def destroy(self) -> None:
        if self._handle:
            LIB.native_fp2_early_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def early_session_create(gate_type, in0, in1, z1, n_copies: int, n_wires: int, values,
                         beta) -> EarlySession:
    """Upload gates, [wire][copy] values and beta once (native_fp2_bridge.py:413-447)."""
    if not early_session_supported():
        raise RuntimeError("truepix HIP early session unavailable: %s" % GPU_FAIL_REASON)
    gt, i0, i1, zr, zi = _gate_arrays(gate_type, in0, in1, z1)
    if isinstance(values, tuple) and len(values) == 2 and isinstance(values[0], np.ndarray):
        vr = np.ascontiguousarray(values[0], dtype=np.uint64)
        vi = np.ascontiguousarray(values[1], dtype=np.uint64)
    else:
        vr, vi = _pack_rows(values, n_wires, n_copies)
    br, bi = (np.ascontiguousarray(x) for x in as_coords(beta))
    handle = LIB.native_fp2_early_create(len(gt), _ptr(gt), _ptr(i0), _ptr(i1), _ptr(zr), _ptr(zi),
                                         n_copies, n_wires, _ptr(vr), _ptr(vi), _ptr(br), _ptr(bi))
    if not handle:
        raise RuntimeError("native_fp2_early_create failed: %s" % last_error())
    return EarlySession(handle, n_wires)


# Gate-type indices (TP/libTruePix/native_fp2_bridge.py:450-472).
GATE_TYPE_IDX = {"mul": 0, "add": 1, "sub": 2, "muxL": 3, "muxR": 4, "or": 5, "xor": 6,
                 "not": 7, "nand": 8, "nor": 9, "nxor": 10, "naab": 11, "pass": 12}


def resolve_gate_type_idx(gate) -> int:
    """Index of a gate prover object: mux by its live mux bit, else gate_type_idx."""
    if getattr(gate, "gate_type", None) == "mux":
        bit = gate.layer.circuit.muxbits[gate.muxbit]
        return 4 if bit else 3
    idx = getattr(gate, "gate_type_idx", None)
    return int(idx) if idx is not None else 12


def cache_gate_tables(gates) -> Tuple[List[int], List[int], List[int]]:
    return ([resolve_gate_type_idx(g) for g in gates], [g.in0 for g in gates],
            [g.in1 for g in gates])


def compute_beta(z: Sequence, init=1) -> Fp2Buf:
    """beta table of length 2^len(z), LSB-first (native_fp2_bridge.py:483-499)."""
    require()
    n = len(z)
    if n == 0:
        return Fp2Buf(np.zeros(0, np.uint64), np.zeros(0, np.uint64), 0)
    zr, zi = (np.ascontiguousarray(x) for x in as_coords(list(z)))
    out_n = 1 << n
    out_r = np.empty(out_n, dtype=np.uint64)
    out_i = np.empty(out_n, dtype=np.uint64)
    ir, ii = (init.real, init.imag) if type(init) is Fp2 else (int(init) % P, 0)
    LIB.native_fp2_compute_beta(_ptr(zr), _ptr(zi), n, ir, ii, _ptr(out_r), _ptr(out_i))
    check("native_fp2_compute_beta")
    return Fp2Buf(out_r, out_i, out_n)


def mle_eval_base(inputs, z: Sequence) -> Fp2:
    """MLE of base-field inputs (zero-padded to 2^len(z)) at the Fp2 point z
    (native_fp2_bridge.py:502-542)."""
    require()
    n_z = len(z)
    n = 1 << n_z
    buf = np.zeros(n, dtype=np.uint64)
    if isinstance(inputs, np.ndarray):
        lim = min(inputs.size, n)
        buf[:lim] = inputs[:lim].astype(np.uint64) % np.uint64(P)
    else:
        lim = min(len(inputs), n)
        vals = []
        for i in range(lim):
            v = inputs[i]
            if type(v) is Fp2:
                if v.imag != 0:
                    raise TypeError("mle_eval_base expects base-field inputs")
                vals.append(v.real)
            else:
                vals.append(int(v) % P)
        buf[:lim] = np.array(vals, dtype=np.uint64)
    zr, zi = (np.ascontiguousarray(x) for x in as_coords(list(z)))
    out_r = np.zeros(1, dtype=np.uint64)
    out_i = np.zeros(1, dtype=np.uint64)
    LIB.native_fp2_mle_eval_base(_ptr(buf), _ptr(zr), _ptr(zi), n_z, _ptr(out_r), _ptr(out_i))
    check("native_fp2_mle_eval_base")
    return _raw(int(out_r[0]), int(out_i[0]))
