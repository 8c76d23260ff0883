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
