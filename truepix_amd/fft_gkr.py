"""fft_gkr on the device (OZ/src/VPD/fft_circuit_GKR.cpp).

The reference is a program, ``./fft_gkr <lg_size> <logfile>``, that the Orion
VPD verifier runs (OZ/src/VPD/vpd_verifier.cpp:80): it builds the IFFT +
evaluation circuit for 2^lg_size coefficients, proves it with its GKR layer
protocols against its own verifier (challenges from glibc rand(), seed 1),
prints "Error, fft gkr failed" when a check fails, and writes
"<v_time> <proof_size> <p_time>" to the log file (:827-845).

``python -m truepix_amd.fft_gkr <lg_size> <logfile>`` is the drop-in: same
arguments, same log line, every layer built and proved by the HIP library
(``tp_fft_gkr`` in include/truepix_hip.h).
"""
from __future__ import annotations

import ctypes
import sys

import numpy as np

from truepix_amd import native_fp2_bridge as nb


def run(lg_size: int):
    """(ok, v_time_s, proof_size, p_time_s) of one fft_gkr run."""
    nb.require()
    v = ctypes.c_double()
    p = ctypes.c_double()
    ps = ctypes.c_int()
    rc = nb.LIB.tp_fft_gkr(int(lg_size), ctypes.byref(v), ctypes.byref(ps), ctypes.byref(p))
    if rc < 0:
        raise RuntimeError("tp_fft_gkr failed: %s" % nb.last_error())
    return bool(rc), v.value, ps.value, p.value


def set_graph(on: bool) -> None:
    """Replay each size's launch sequence from a captured hipGraph (default),
    or enqueue every launch directly."""
    nb.require()
    nb.LIB.tp_fft_gkr_set_graph(1 if on else 0)


def trace(lg_size: int, layers: bool = False) -> dict:
    """Everything one run computes: round polynomials (n x 6 limbs), the claim
    after every layer with the final alpha / beta (7 x 2 limbs), optionally
    the last IFFT layer, the scaled layer and the 64 sums, the proof size and
    the pass flag."""
    nb.require()
    lg = int(lg_size)
    n = 1 << lg
    n_rounds = (lg + 6) + lg + 2 * lg * lg
    polys = np.zeros(6 * n_rounds, np.uint64)
    claims = np.zeros(14, np.uint64)
    lay = np.zeros(4 * n + 128, np.uint64) if layers else None
    npol = ctypes.c_int()
    ps = ctypes.c_int()
    u64p = ctypes.POINTER(ctypes.c_uint64)
    rc = nb.LIB.tp_fft_gkr_trace(lg, polys.ctypes.data_as(u64p), n_rounds, ctypes.byref(npol),
                                 claims.ctypes.data_as(u64p),
                                 lay.ctypes.data_as(u64p) if layers else None, ctypes.byref(ps))
    if rc < 0:
        raise RuntimeError("tp_fft_gkr_trace failed: %s" % nb.last_error())
    return {"ok": rc, "polys": polys, "n_polys": npol.value, "claims": claims, "layers": lay,
            "proof_size": ps.value}


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 2:
        print("usage: python -m truepix_amd.fft_gkr <lg_size> <logfile>", file=sys.stderr)
        return 2
    ok, v_time, proof_size, p_time = run(int(argv[0]))
    if not ok:
        print("Error, fft gkr failed", file=sys.stderr)
    with open(argv[1], "w") as fh:
        fh.write("%f %d %f\n" % (v_time, proof_size, p_time))
    return 0


if __name__ == "__main__":
    sys.exit(main())
