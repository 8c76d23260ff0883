"""CPU checks of the C ABI: the library loads and exports every declared symbol,
and without a GPU the bridge fails loudly instead of computing on the host."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "truepix_amd", "lib", "libtruepix_hip.so")


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b((?:native_fp2|tp)_[A-Za-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_drop_in_set():
    names = set(declared_symbols())
    drop_in = {"native_fp2_init", "native_fp2_cuda_available", "native_fp2_fold",
               "native_fp2_expand", "native_fp2_sumcheck_early", "native_fp2_compute_beta",
               "native_fp2_mle_eval_base", "native_fp2_early_create", "native_fp2_early_claim",
               "native_fp2_early_fold", "native_fp2_early_finish", "native_fp2_early_n_copies",
               "native_fp2_early_n_wires", "native_fp2_early_destroy"}
    assert drop_in <= names


def test_library_exports_every_declared_symbol():
    assert os.path.isfile(LIB), "build first: python -m truepix_amd.build"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_bridge_binds_all_symbols():
    from truepix_amd import native_fp2_bridge as nb
    lib = nb.load_library()
    for n in declared_symbols():
        assert getattr(lib, n).argtypes is not None or n in (
            "tp_last_error",), n


def test_no_host_fallback_without_device():
    from truepix_amd import native_fp2_bridge as nb
    if nb.AVAILABLE:
        pytest.skip("a GPU is present; the no-device path is not reachable here")
    assert not nb.GPU_ACTIVE
    with pytest.raises(RuntimeError):
        nb.fold_values([1, 2, 3, 4], 5)
    with pytest.raises(RuntimeError):
        nb.compute_beta([1, 2])
    with pytest.raises(RuntimeError):
        nb.early_session_create([12], [0], [0], [1], 2, 1, [[1, 2]], [1, 1])
    # the raw entry points report the missing device instead of computing
    lib = nb.load_library()
    assert lib.native_fp2_cuda_available() == 0
    assert lib.native_fp2_init() == -1
    assert lib.native_fp2_early_claim(None, (ctypes.c_uint64 * 8)()) == -1


def _integration_snippet():
    """The `_load()` replacement block of INTEGRATION.md, as a maintainer would
    paste it into TP/libTruePix/native_fp2_bridge.py."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    start = text.index("# TP/libTruePix/native_fp2_bridge.py, _load() replacement")
    return text[start:text.index("```", start)]


def test_integration_binding_raises_on_device_errors(monkeypatch):
    """Runs the documented binding against the library: with no device, the
    void entry points the reference's callers never check raise instead of
    leaving their outputs unwritten; a call that needs no device does not."""
    import numpy as np
    from truepix_amd import native_fp2_bridge as nb
    monkeypatch.setenv("TRUEPIX_HIP_LIB", LIB)
    ns = {"os": os, "ctypes": ctypes, "_LIB": None, "_GPU_LIB": None, "AVAILABLE": False,
          "GPU_ACTIVE": False, "GPU_FAIL_REASON": "",
          # the reference's own binders set the CUDA library's argtypes; the
          # drop-in symbols have the same signatures here
          "_bind_batch": nb.bind, "_bind_early": lambda lib: None}
    exec(compile(_integration_snippet(), "INTEGRATION.md", "exec"), ns)
    ns["_load"]()
    lib = ns["_HIP"]
    assert lib is not None
    u = lambda n: np.zeros(n, np.uint64)
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    ir, ii, outr, outi = u(8), u(8), u(4), u(4)
    lib.native_fp2_fold(p(ir), p(ii), 0, 1, 0, p(outr), p(outi))     # nothing to do: no error
    z = u(41)
    with pytest.raises(RuntimeError, match="n_z=41 too large"):
        lib.native_fp2_compute_beta(p(z), p(z), 41, 1, 0, p(u(2)), p(u(2)))
    if ns["AVAILABLE"]:
        lib.native_fp2_fold(p(ir), p(ii), 4, 1, 0, p(outr), p(outi))
        return
    assert "HIP device unavailable" in ns["GPU_FAIL_REASON"] or "no HIP device" in ns["GPU_FAIL_REASON"]
    calls = [
        lambda: lib.native_fp2_fold(p(ir), p(ii), 4, 1, 0, p(outr), p(outi)),
        lambda: lib.native_fp2_expand(p(ir), p(ii), 4, 2, p(u(8)), p(u(8))),
        lambda: lib.native_fp2_compute_beta(p(z), p(z), 3, 1, 0, p(u(8)), p(u(8))),
        lambda: lib.native_fp2_mle_eval_base(p(u(8)), p(z), p(z), 3, p(u(1)), p(u(1))),
    ]
    for call in calls:
        with pytest.raises(RuntimeError, match="HIP device unavailable"):
            call()
    g = np.zeros(1, np.int32)
    gi = g.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    with pytest.raises(RuntimeError, match="HIP device unavailable"):
        lib.native_fp2_sumcheck_early(1, gi, gi, gi, p(u(1)), p(u(1)), 2, 1, *([p(u(2))] * 12),
                                      p(u(8)))
