"""Host wrapper of the degree-2 sumcheck prover (tp_sc2_prove).

The prover follows OZ/src/VPD/fft_circuit_GKR.cpp:171-209 (round polynomial
(a, b, c) of sum_i A_i(X) V_i(X), fold with the round challenge) over two
resident device tables; challenges come from the TruePix transcript seeded by
put(["c2", seed]).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from truepix_amd import native_fp2_bridge as nb

P = (1 << 61) - 1


@dataclass
class Deg2Proof:
    polys: np.ndarray    # [nv, 6]: a.re a.im b.re b.im c.re c.im
    chals: np.ndarray    # [nv, 2]
    final: np.ndarray    # [4]: f~(r), g~(r)


def synthetic_tables(nv: int, seed: int):
    """Two tables of 2^nv uniform canonical Fp2 values (interleaved re, im),
    drawn from numpy PCG64(seed) with rejection of the single value 2^61 - 1."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(2):
        n = 2 << nv
        x = rng.integers(0, 1 << 61, size=n, dtype=np.uint64)
        bad = x == P
        while bad.any():
            x[bad] = rng.integers(0, 1 << 61, size=int(bad.sum()), dtype=np.uint64)
            bad = x == P
        out.append(x)
    return out[0], out[1]


def _u64p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def prove_deg2_device(d_f: int, d_g: int, nv: int, seed: int = 0, fixed=None,
                      stream: int = 0) -> Deg2Proof:
    """Prove over device tables given as raw device addresses (16-B aligned)."""
    nb.require()
    polys = np.zeros((nv, 6), dtype=np.uint64)
    chals = np.zeros((nv, 2), dtype=np.uint64)
    final = np.zeros(4, dtype=np.uint64)
    fx = None
    if fixed is not None:
        fx = np.ascontiguousarray(fixed, dtype=np.uint64).reshape(-1)
        if fx.size != 2 * nv:
            raise ValueError("fixed challenges need 2*nv limbs")
    rc = nb.LIB.tp_sc2_prove(ctypes.c_void_p(d_f), ctypes.c_void_p(d_g), nv, seed,
                             _u64p(fx) if fx is not None else None, _u64p(polys), _u64p(chals),
                             _u64p(final), ctypes.c_void_p(stream) if stream else None)
    if rc != 0:
        raise RuntimeError("tp_sc2_prove failed: %s" % nb.last_error())
    return Deg2Proof(polys, chals, final)


def to_device(table: np.ndarray):
    """Copy an interleaved uint64 table into a torch device tensor (int64 view)."""
    import torch
    t = torch.from_numpy(np.ascontiguousarray(table).view(np.int64)).to("cuda")
    torch.cuda.synchronize()
    return t


def prove_deg2_host(f: np.ndarray, g: np.ndarray, seed: int = 0, fixed=None) -> Deg2Proof:
    """Upload host tables, prove, return the proof (tables are not modified)."""
    nv = (f.size // 2).bit_length() - 1
    if f.size != 2 << nv or g.size != f.size:
        raise ValueError("tables must hold 2^nv Fp2 values each")
    df, dg = to_device(f), to_device(g)
    return prove_deg2_device(df.data_ptr(), dg.data_ptr(), nv, seed, fixed)


def set_timing(on: bool) -> None:
    nb.LIB.tp_sc2_set_timing(1 if on else 0)


def set_ahead(on: bool) -> None:
    """Queue an untimed prove's launches ahead of their transcript steps, or
    make each after the step before it (default); same proof either way."""
    nb.LIB.tp_sc2_set_ahead(1 if on else 0)


def set_block_pass(on: bool) -> None:
    """Rounds 0 to 3 from the input's block products and one pass folding the
    input into T_4 (default, for 2^14 to 2^28 entries), or rounds 0 and 1 from
    its quad products and passes folding by two challenges; same proof either
    way."""
    nb.LIB.tp_sc2_set_block_pass(1 if on else 0)


def stats():
    ms, by, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    nb.LIB.tp_sc2_stats(ctypes.byref(ms), ctypes.byref(by), ctypes.byref(n))
    return ms.value, by.value, n.value


def tail_stats():
    """(ms, launches) of the one-workgroup tail rounds of the last prove."""
    ms, n = ctypes.c_double(), ctypes.c_int()
    nb.LIB.tp_sc2_tail_stats(ctypes.byref(ms), ctypes.byref(n))
    return ms.value, n.value


CLASSES = ("first", "fold01", "fold", "tail", "mid", "pair")


def class_stats():
    """Per launch class of the last prove (timing on): {"first" | "fold01" |
    "fold" | "tail" | "mid" | "pair": {"ms", "bytes", "launches"}}.  "first"
    is the pass over the input giving rounds 0 and 1, "fold01" round 2
    folding the input by r_0 and r_1 (and giving round 3's sums too on inputs
    above 2^22 entries), "fold" the streaming fold+round launches, "tail" the
    one-workgroup rounds, "mid" the rounds on per-workgroup slices, "pair"
    the later passes folding by two challenges."""
    out = (ctypes.c_double * 18)()
    nb.LIB.tp_sc2_class_stats(out, 18)
    return {c: {"ms": out[3 * i], "bytes": out[3 * i + 1], "launches": int(out[3 * i + 2])}
            for i, c in enumerate(CLASSES)}


def tail_profile():
    """Per-exchange latency of the one-workgroup tail in the last timed
    prove (an exchange is one host round trip for one or two rounds):
    {"rounds" (exchanges), "device_us" (an exchange's fold and sums),
    "wait_us" (waiting for the host's challenges), "host_us" (the host's
    transcript steps), "round_device_us" (per exchange)}."""
    out = (ctypes.c_double * 44)()
    nb.LIB.tp_sc2_tail_profile(out, 44)
    n = int(out[0])
    return {"rounds": n, "device_us": round(out[1], 3), "wait_us": round(out[2], 3),
            "host_us": round(out[3], 3), "round_device_us": [round(out[4 + i], 2) for i in range(n)]}


def mid_profile():
    """Per-exchange latency of the sliced mid rounds in the last timed prove
    (an exchange is one host round trip for one or two rounds):
    {"rounds" (exchanges), "host_us" (last partials to the challenges at
    workgroup 0), "relay_us" (to the last workgroup), "device_us" (challenges
    to the next partials), "spread_us" (first to last workgroup's partials)}."""
    out = (ctypes.c_double * 5)()
    nb.LIB.tp_sc2_mid_profile(out, 5)
    return {"rounds": int(out[0]), "host_us": round(out[1], 3), "relay_us": round(out[2], 3),
            "device_us": round(out[3], 3), "spread_us": round(out[4], 3)}


def schedule():
    """(mid_groups, tail_in, mid_end): workgroups of the sliced mid rounds,
    the largest input (entries per table) of the one-workgroup tail, and the
    input the mid rounds hand to the tail."""
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    nb.LIB.tp_sc2_schedule(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


def set_gate_ticks(ticks: int) -> None:
    """Bound of the mid launch's arrival gate in wall-clock ticks (negative:
    the default 20 ms; 0: give way unless every workgroup already arrived)."""
    nb.LIB.tp_sc2_set_gate_ticks(int(ticks))


def set_spin_ticks(ticks: int) -> None:
    """Bound of every device-side wait of a prove in wall-clock ticks
    (negative: the default 10 s); a small bound makes a prove fail with
    every kernel leaving on its own."""
    nb.LIB.tp_sc2_set_spin_ticks(int(ticks))


def mid_fallbacks() -> int:
    """Mid launches of the last prove that gave way to streaming launches."""
    return int(nb.LIB.tp_sc2_mid_fallbacks())
