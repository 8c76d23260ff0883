#!/usr/bin/env python3
"""Benchmark: degree-2 sumcheck prover over two resident 2^24-entry Fp2 tables
(24 rounds, TruePix transcript challenges) on one MI355X.

One step = one full 24-round prove (round sums, fused folds, host transcript)
from tables already resident in HBM.  Prints ONE JSON line:
  value      whole-job fold HBM throughput: the algorithmic bytes of a prove,
             48 (2^25 - 2) B (round r reads 2 x 2^(24-r) x 16 B and writes
             half of that, whatever the schedule), / wall time per step;
  roofline   the launch class moving the most HBM bytes per prove (of the
             first pass over the input, the round-2 fold from the input and the
             streaming fold+round launches): the bytes its launches read and
             write / its summed HIP-event time (events on the stream it runs
             on), against 8 TB/s; every class is listed beside it;
  cpu_baseline  the CPU oracle (oracle/fp2_oracle.c, a port) proving the same
             2 x 2^24 workload single-threaded on this host, repeated for at
             least 10 s, in GB/s of the same kind;
  rs_encode  the Reed-Solomon NTT, 64 columns x 2^20 -> 2^22 (side measurement).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GKR frames proved/s + fold HBM GB/s; batched/served vs plain; p50/p99 latency"
PEAK_HBM_GBS = 8000.0
NV = 24


def traffic_file(kernel: str) -> str:
    return os.path.join(ROOT, "profiles", "%s_traffic.json" % kernel.replace("k_", "", 1))


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (each scaled by the calibration copy), or
    None when the summary covers another kernel."""
    try:
        with open(traffic_file(kernel)) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        return None
    if rec.get("kernel") != kernel:
        return None
    return rec.get("bytes_per_launch")


def measured_valu(kernel: str):
    """The committed VALU counter summary of `kernel` (rocprofv3 SQ_INSTS_VALU
    and the kernel's loop instruction counts), or None when there is none."""
    path = os.path.join(ROOT, "profiles", "%s_valu.json" % kernel.replace("k_", "", 1))
    try:
        with open(path) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        return None
    if rec.get("kernel") != kernel:
        return None
    keep = ("sq_insts_valu_per_launch", "valu_busy_at_4_cycles", "loop_valu_insts_per_wave_iteration",
            "loop_mad_u64_u32_per_wave_iteration", "note")
    out = {k: rec[k] for k in keep if k in rec}
    out["source"] = os.path.relpath(path, ROOT)
    return out


def alg_bytes(nv: int) -> float:
    """Algorithmic HBM bytes of one nv-round prove (both tables, 16 B per Fp2):
    round r reads the 2 x 2^(nv-r) entries of its tables and writes the
    2 x 2^(nv-r-1) folded ones, 48 x 2^(nv-r) B, summed over the rounds."""
    return 48.0 * ((1 << (nv + 1)) - 2)


SC2_KERNELS = {"first": "k_sc2_blocks_mfma", "fold01": "k_sc2_fold16", "pair": "k_sc2_fold2_pairs_cached",
               "fold": "k_sc2_fold_round", "tail": "k_sc2_tail_host", "mid": "k_sc2_mid"}


# 32 x 32-bit multiply lane-ops of one Fp2 twiddle product (3 Karatsuba Fp
# products of 4 limb products each), and MI355X's issue rate for them.
MUL_LANE_OPS_PER_BUTTERFLY = 12
INT_MUL_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12


def rs_counters():
    """Per-encode counter figures of the RS encode committed under profiles/
    (rocprofv3 --pmc runs of this bench): VALU lane-instructions per
    butterfly and HBM bytes per encode; None where absent."""
    try:
        with open(os.path.join(ROOT, "profiles", "rs_encode_64x2e20_counters.json")) as fh:
            rec = json.load(fh)
        return rec.get("valu_lane_insts_per_butterfly"), rec.get("hbm_bytes_per_encode")
    except (OSError, ValueError):
        return None, None


def bench_rs_encode(steps: int = 5, cols: int = 64, lc: int = 20, lo: int = 22, cpu: bool = True):
    """Reed-Solomon encode of 64 columns x 2^20 -> 2^22 (resident in HBM)."""
    import torch
    from truepix_amd import ntt
    from truepix_amd.sumcheck import synthetic_tables
    coef, _ = synthetic_tables(lc + 6, seed=33)          # 2^26 = 64 x 2^20 Fp2
    dc = torch.from_numpy(coef.view("int64")).to("cuda")
    do = torch.empty(cols * (2 << lo), dtype=torch.int64, device="cuda")
    st = torch.cuda.Stream()                             # a real stream: events and kernels share it
    ntt.rs_encode_device(dc.data_ptr(), cols, lc, lo, do.data_ptr(), stream=st.cuda_stream)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(st)
    for _ in range(steps):
        ntt.rs_encode_device(dc.data_ptr(), cols, lc, lo, do.data_ptr(), stream=st.cuda_stream)
    ev1.record(st)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    alg = cols * ((16 << lc) + (16 << lo))
    # the reference's network: replicate to 2^lo, then lc radix-2 stages of
    # 2^(lo-1) butterflies, each one Fp2 twiddle product (RS_polynomial.cpp:46-130)
    bfly = cols * lc * (1 << (lo - 1))
    ops = bfly * MUL_LANE_OPS_PER_BUTTERFLY
    coef_host = coef[: 2 << lc].copy()
    del dc, do, coef
    out = {"workload": "RS encode 64 x 2^20 -> 2^22 Fp2 (rate 1/4)", "ms": round(ms, 3),
           "alg_GBps": round(alg / ms / 1e6, 1), "Gbutterflies_per_s": round(bfly / ms / 1e6, 2),
           "alg_bytes": alg,
           "roofline": {"bound": "valu", "achieved": round(ops / ms / 1e9, 3),
                        "peak": INT_MUL_PEAK_TOPS, "unit": "T int32-multiply lane-ops/s",
                        "frac": round(ops / ms / 1e9 / INT_MUL_PEAK_TOPS, 4),
                        "traffic": rs_counters()[1],
                        "hbm_frac": round(alg / ms / 1e6 / PEAK_HBM_GBS, 4),
                        "valu_lane_insts_per_butterfly": rs_counters()[0],
                        "note": "useful work = 12 32x32-bit multiply lane-ops per butterfly "
                                "(one Fp2 twiddle product: 3 Fp products by Karatsuba, 4 limb "
                                "products each) of the reference's radix-2 network; peak = "
                                "256 CUs x 64 lanes x 2.4 GHz (the published 157.3 TFLOP/s "
                                "fp32 counts packed FMAs: 4 flops per lane-cycle, the integer "
                                "multiply-add has no packed form); traffic: HBM bytes of one "
                                "encode (FETCH_SIZE + WRITE_SIZE, scaled), including the "
                                "intermediate between the two passes; "
                                "valu_lane_insts_per_butterfly: SQ_INSTS_VALU x 64 of both "
                                "passes / butterflies (profiles/rs_encode_64x2e20_counters.json)"}}
    if cpu:
        out["cpu_baseline"] = cpu_rs_baseline(coef_host, lc, lo, ms, cols)
    return out


def cpu_rs_baseline(col, lc: int, lo: int, gpu_ms: float, cols: int, seconds: float = 6.0):
    """One column of the same encode on the host: the reference's own
    fast_fourier_transform (RS_polynomial.cpp compiled from source,
    oracle/_ref) on one core, and the oracle's coset DFTs on every core of
    this process's share, one column per thread."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    from oracle.oracle import Oracle, RefRS
    col_bytes = (16 << lc) + (16 << lo)
    rec = {}
    if RefRS.available():
        impl, kind, what = RefRS(), "reference", "oracle/_ref fast_fourier_transform (the reference's code)"
    else:
        impl, kind, what = Oracle(), "port", "oracle/fp2_oracle.c or_fft"
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds / 2 or n == 0:
        impl.fft(col, lc, lo)
        n += 1
    dt = (time.perf_counter() - t0) / n
    rec.update({"value": round(col_bytes / dt / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": kind,
                "sample": "%d column(s) of 2^%d -> 2^%d, %s, %.2f s a column" % (n, lc, lo, what, dt),
                "ms_per_64_columns": round(dt * cols * 1e3, 1),
                "gpu_speedup": round(dt * cols * 1e3 / gpu_ms, 1)})
    threads = host_threads()
    orc = Oracle()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda _: orc.fft(col, lc, lo), range(threads)))
    dt = time.perf_counter() - t0
    rec["all_cores"] = {"value": round(threads * col_bytes / dt / 1e9, 4), "unit": "GB/s",
                        "cores": threads, "kind": "port",
                        "sample": "%d columns at once, one per thread, oracle/fp2_oracle.c or_fft "
                                  "(coset DFTs), %.2f s" % (threads, dt),
                        "ms_per_64_columns": round(dt * cols / threads * 1e3, 1)}
    return rec


def host_threads() -> int:
    """Host cores this process may use (its affinity, capped by OMP_NUM_THREADS)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def bench_gkr_frame(frames: int = 5, circuit: str = "gray", n_copies: int = 1350):
    """Plain GKR prove of one frame (C1 shape: gray, 1350 macroblocks, 384
    samples each) with the Python-driven device prover (gkr.CircuitProverNIZK:
    protocol steps in Python, kernels through the ABI); proof bytes as the
    reference's."""
    import tempfile
    from truepix_amd import circuits, gkr
    from truepix_amd.fp2_link import full_input_rows, make_manifest, random_private_inputs, write_manifest
    from truepix_amd.vole import ExtVOLEGenerator
    pid = 4242
    work = tempfile.mkdtemp(prefix="tpbench_")
    gen = ExtVOLEGenerator(seed=777)
    gen.generate_deltas()
    gen.generate_vole_instances(500)
    gen.save_to_files(os.path.join(work, "vole_ext_prover_%d.json" % pid),
                      os.path.join(work, "vole_ext_verifier_%d.json" % pid))
    write_manifest(os.path.join(work, "orion_commit.json"), make_manifest(n_copies, seed=1))
    parsed = circuits.circuit(circuit)
    rows = full_input_rows(parsed, random_private_inputs(circuit, parsed, n_copies, seed=105))
    times = []
    proof = None
    for i in range(frames + 1):
        t0 = time.perf_counter()
        _, prv = gkr.ProverClass.from_pws(parsed, n_copies)
        prv.pid = pid
        prv.vole_dir = work
        prv.manifest_path = os.path.join(work, "orion_commit.json")
        proof = prv.run(rows)
        if i:                                  # the first frame is warm-up
            times.append(time.perf_counter() - t0)
    times.sort()
    med = times[len(times) // 2]
    return {"workload": "plain GKR prove, %s circuit, %d copies x 384 samples, 1 frame "
                        "(proof bytes identical to the reference's)" % (circuit, n_copies),
            "frames_per_s": round(1.0 / med, 3), "ms_per_frame_median": round(med * 1e3, 2),
            "ms_per_frame_min": round(times[0] * 1e3, 2), "proof_bytes": len(proof),
            "phases_ms": {k: round(v * 1e3, 2) for k, v in getattr(prv, "timings", {}).items()}}


# --- whole frames (C1 plain, C4 batched, C5 served) through the device frame prover

FRAME_SIZES = {"64x64": 16, "240p": 300, "480p": 1350, "720p": 3600, "1080p": 8040}
CIRCUITS = ("gray", "crop", "mask", "invert")


class _FrameFactory:
    """Synthetic frames: uint8 macroblock samples from PCG64(seed) (crop adds
    its keep bit), a fresh pid, VOLE correlations and commitment binding per
    frame (VOLE U, V uniform nonzero from PCG64; the prover only reads U, V)."""

    def __init__(self, seed: int):
        import numpy as np
        from truepix_amd import circuits, frames
        self.np = np
        self.rng = np.random.Generator(np.random.PCG64(seed))
        self.circ = {n: frames.FrameCircuit(circuits.circuit(n)) for n in CIRCUITS}
        self.inputs = {}

    def inputs_for(self, name: str, n_copies: int, variant: int = 0):
        key = (name, n_copies, variant)
        if key not in self.inputs:
            from truepix_amd.fp2_link import crop_bits
            c = self.circ[name]
            arr = self.rng.integers(0, 256, size=(n_copies, c.n_private), dtype=self.np.uint8)
            if name == "crop":
                arr[:, -1] = crop_bits(n_copies, 0.2, 0.8, 0.2, 0.8).astype(self.np.uint8)
            self.inputs[key] = arr
        return self.inputs[key]

    def job(self, name: str, n_copies: int, k: int, variant: int = 0):
        from truepix_amd.fp2_link import commit_binding, make_manifest
        P = (1 << 61) - 1
        vole = tuple(self.rng.integers(1, P, size=500, dtype=self.np.int64).astype(self.np.uint64)
                     for _ in range(4))
        binding = commit_binding(make_manifest(n_copies, seed=k))
        return (self.circ[name], self.inputs_for(name, n_copies, variant), 20000 + k, vole, binding)


def bench_frames_plain(frames_n: int = 12, circuit: str = "gray", n_copies: int = 1350):
    """C1: plain prove of one frame at a time (gray, 1350 macroblocks x 384
    samples) by the device frame prover, one worker."""
    from truepix_amd import frames
    fac = _FrameFactory(105)
    times, phases = [], []
    with frames.FramePool(1) as pool:
        for i in range(frames_n + 2):
            spec = fac.job(circuit, n_copies, i)
            t0 = time.perf_counter()
            res = pool.result(pool.submit(*spec))
            if i >= 2:                            # two warm-up frames
                times.append(time.perf_counter() - t0)
                phases.append(res.phases_ms)
    times.sort()
    med = times[len(times) // 2]
    ph = {k: round(sorted(p[k] for p in phases)[len(phases) // 2], 3) for k in phases[0]}
    return {"workload": "plain GKR prove, %s circuit, %d copies x 384 samples, one frame at a time "
                        "(device frame prover, proof bytes as the reference's)" % (circuit, n_copies),
            "frames_per_s": round(1.0 / med, 2), "ms_per_frame_median": round(med * 1e3, 3),
            "ms_per_frame_min": round(times[0] * 1e3, 3), "proof_bytes": len(res.proof),
            "phases_ms_median": ph}


def cpu_frame_baseline(n_copies: int = 1350, seconds: float = 6.0, seed: int = 105):
    """The arithmetic of one plain gray frame on the host: the reference's own
    CPU library (TP/native_fp2/native_fp2.cpp compiled from source, oracle/_ref;
    the oracle's restatement where that is absent) running the frame's early
    session (the 11 claim + fold rounds over 512 wires x 2048 copies and 384
    gates), its input MLE over 2^20 inputs, and CPython's SHA-256 over the
    8 MiB of outputs the transcript absorbs.  The late rounds, the VOLE
    commitments and the transcript framing (under 1 ms together on the
    device path) are not counted, so this is a lower bound on the CPU's time
    per frame."""
    import ctypes
    import hashlib
    import numpy as np
    from oracle.oracle import Oracle, RefNative, _p
    from truepix_amd import circuits
    from truepix_amd.fp2_link import full_input_rows, random_private_inputs
    from truepix_amd.gkr import parse_layers
    P = (1 << 61) - 1
    rng = np.random.Generator(np.random.PCG64(seed))
    parsed = circuits.circuit("gray")
    _, (lay,) = parse_layers(parsed)
    rows = full_input_rows(parsed, random_private_inputs("gray", parsed, n_copies, seed=seed))
    cp = 1 << (n_copies - 1).bit_length()
    nw = 1 << lay.in_bits
    vr = np.zeros((nw, cp), np.uint64)
    vr[:rows.shape[1], :n_copies] = rows.T.astype(np.uint64)
    vi = np.zeros_like(vr)
    gt, i0, i1 = (np.ascontiguousarray(a, np.int32) for a in (lay.type_idx([0]), lay.in0, lay.in1))
    z = rng.integers(0, P, size=(4, lay.n_gates), dtype=np.int64).astype(np.uint64)
    beta = rng.integers(0, P, size=(2, cp), dtype=np.int64).astype(np.uint64)
    chal = rng.integers(0, P, size=(64, 2), dtype=np.int64).astype(np.uint64)
    inputs = np.zeros(1 << 20, np.uint64)
    inputs[: rows.size] = rows.reshape(-1).astype(np.uint64)
    zin = rng.integers(0, P, size=(2, 20), dtype=np.int64).astype(np.uint64)
    outputs = np.zeros(cp * 512, np.uint64).tobytes()
    ref = RefNative.available()
    impl = RefNative() if ref else None
    orc = Oracle()
    out8 = (ctypes.c_uint64 * 8)()
    o2 = (ctypes.c_uint64 * 1)(), (ctypes.c_uint64 * 1)()

    def one_frame():
        hashlib.sha256(outputs).digest()
        if ref:
            h = impl.early_create(lay.n_gates, _p(gt), _p(i0), _p(i1), _p(z[0]), _p(z[1]), cp, nw,
                                  _p(vr), _p(vi), _p(beta[0]), _p(beta[1]))
            for r in range(cp.bit_length() - 1):
                impl.early_claim(h, out8)
                impl.early_fold(h, int(chal[r, 0]), int(chal[r, 1]))
            impl.early_destroy(h)
            impl.mle_eval_base(_p(inputs), _p(zin[0]), _p(zin[1]), 20, o2[0], o2[1])
        else:
            wr, wi, br_, bi_ = vr.copy(), vi.copy(), beta[0].copy(), beta[1].copy()
            n = cp
            for r in range(cp.bit_length() - 1):
                orc.early_claim(lay.n_gates, _p(gt), _p(i0), _p(i1), _p(z[0]), _p(z[1]), n,
                                _p(wr), _p(wi), _p(br_), _p(bi_), out8)
                nr, ni = np.zeros(nw * n // 2, np.uint64), np.zeros(nw * n // 2, np.uint64)
                nbr, nbi = np.zeros(n // 2, np.uint64), np.zeros(n // 2, np.uint64)
                orc.early_fold(nw, n, _p(wr), _p(wi), _p(br_), _p(bi_), int(chal[r, 0]),
                               int(chal[r, 1]), _p(nr), _p(ni), _p(nbr), _p(nbi))
                wr, wi, br_, bi_, n = nr, ni, nbr, nbi, n // 2
            orc.mle_eval_base(_p(inputs), _p(zin[0]), _p(zin[1]), 20, o2[0], o2[1])
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < seconds or k == 0:
        one_frame()
        k += 1
    dt = (time.perf_counter() - t0) / k
    return {"value": round(1.0 / dt, 3), "unit": "frames/s", "cores": 1,
            "kind": "reference" if ref else "port", "ms_per_frame": round(dt * 1e3, 1),
            "sample": "%d frame(s) of gray x %d: early session (11 claim + fold rounds over 512 x "
                      "2048), input MLE over 2^20, SHA-256 of the 8 MiB outputs, by %s; late "
                      "rounds, VOLE and transcript framing not counted (a lower bound)"
                      % (k, n_copies, "the reference's native_fp2.cpp compiled from source "
                         "(oracle/_ref)" if ref else "oracle/fp2_oracle.c")}


# C4 pools: stream-W = W stream workers (stream-1 is the plain prover, one
# frame after another); lockstep-DxB = D lockstep batch drivers of up to B
# frames per launch set.
BATCHED_CONFIGS = [("stream-1", dict(n_workers=1)), ("stream-8", dict(n_workers=8)),
                   ("stream-16", dict(n_workers=16)),
                   ("lockstep-1x256", dict(n_workers=1, batched=True, max_batch=256)),
                   ("lockstep-2x128", dict(n_workers=2, batched=True, max_batch=128))]


def bench_batched(n_frames: int = 256, seed: int = 4, variants=None):
    """C4: 256 frames, circuit uniform over gray/crop/mask/invert, size uniform
    over 64x64 / 240p / 480p / 720p; all queued at once.  Pools: stream workers
    (W host threads, one HIP stream each, one frame at a time; W = 1 is the
    plain prover one frame after another) and lockstep batch drivers (D
    drivers, up to B frames per launch set, one launch per protocol step
    across the batch)."""
    from truepix_amd import frames
    import numpy as np
    if variants is None:
        variants = BATCHED_CONFIGS
    fac = _FrameFactory(seed)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    sizes = ["64x64", "240p", "480p", "720p"]
    plan = [(CIRCUITS[rng.integers(0, 4)], FRAME_SIZES[sizes[rng.integers(0, 4)]]) for _ in range(n_frames)]
    specs = [fac.job(c, n, k, variant=k % 4) for k, (c, n) in enumerate(plan)]
    out = {"workload": "%d frames, circuits uniform over %s, sizes uniform over %s macroblocks"
                       % (n_frames, "/".join(CIRCUITS), "/".join(str(FRAME_SIZES[s]) for s in sizes)),
           "total_copies": int(sum(n for _, n in plan)), "by_config": {}}
    ref_proofs = None
    for label, kw in variants:
        with frames.FramePool(**kw) as pool:
            if kw.get("batched"):
                # sizes every driver's arena and lets the schedule tuner settle
                # (three claim sizes, then three fold sizes, one batch each)
                for _ in range(6):
                    for j in pool.submit_many(specs):
                        pool.result(j)
                warm = []
            else:
                warm = [pool.submit(*specs[k]) for k in range(min(2 * kw["n_workers"], n_frames))]
            for j in warm:
                pool.result(j)
            m0 = pool.memory() if kw.get("batched") else None
            pf0 = pool.batch_profile() if kw.get("batched") else None
            t0 = time.perf_counter()
            jobs = pool.submit_many(specs)
            res = [pool.result(j) for j in jobs]
            wall = time.perf_counter() - t0
            m1 = pool.memory() if kw.get("batched") else None
            pf1 = pool.batch_profile() if kw.get("batched") else None
            pool_tuning = pool.tuning() if kw.get("batched") else None
        proofs = [r.proof for r in res]
        if ref_proofs is None:
            ref_proofs = proofs
        # proved/s on the pool's clock: from the submission to the last proof done
        # (the wall figure adds copying every proof into Python)
        dt = max(r.t_end for r in res) - min(r.t_submit for r in res)
        rec = {"frames_per_s": round(n_frames / dt, 2), "seconds": round(dt, 4),
               "wall_frames_per_s": round(n_frames / wall, 2),
               "proofs_equal_plain": all(a == b for a, b in zip(ref_proofs, proofs))}
        if m1 is not None:
            nb_ = m1["batches"] - m0["batches"]
            rec["batches"] = nb_
            rec["mean_batch"] = round((m1["frames"] - m0["frames"]) / max(1, nb_), 1)
            rec["peak_admitted_GB"] = round(m1["peak"] / 1e9, 2)
            ph = {}
            for r in res:
                for k, v in r.phases_ms.items():
                    ph[k] = max(ph.get(k, 0.0), v)
            rec["batch_phases_ms_max"] = {k: round(v, 2) for k, v in ph.items()}
            rec["batch_profile"] = {k: round(pf1[k] - pf0[k], 2) for k in pf1}
            rec["schedule_tuner"] = pool_tuning
        out["by_config"][label] = rec
        print("bench: batched %s %.1f frames/s" % (label, rec["frames_per_s"]), file=sys.stderr,
              flush=True)
    base = out["by_config"][variants[0][0]]["frames_per_s"]
    best = max(out["by_config"], key=lambda k: out["by_config"][k]["frames_per_s"])
    out["best_config"] = best
    out["frames_per_s"] = out["by_config"][best]["frames_per_s"]
    out["vs_plain"] = round(out["frames_per_s"] / base, 2)
    return out


def bench_served(n_jobs: int = 2000, rate: float = 200.0, seed: int = 5, variants=None):
    """C5: Poisson arrivals (rate jobs/s, PCG64(seed) exponential gaps) of
    gray/crop/mask/invert jobs over 240p / 480p / 720p / 1080p frames; each job
    is submitted at its arrival time and admitted as soon as a prover frees up
    (stream workers) or at the next launch set of a lockstep driver (jobs that
    arrive while a batch runs form the next batch, within the device budget).
    Latency = proof done - arrival.  The plain figure is each job's latency
    alone on an idle prover."""
    from truepix_amd import frames
    import numpy as np
    if variants is None:
        variants = [("stream-8", dict(n_workers=8)),
                   ("lockstep-2x64", dict(n_workers=2, batched=True, max_batch=64))]
    fac = _FrameFactory(seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = ["240p", "480p", "720p", "1080p"]
    gaps = rng.exponential(1.0 / rate, size=n_jobs)
    arrive = np.cumsum(gaps)
    plan = [(CIRCUITS[rng.integers(0, 4)], FRAME_SIZES[sizes[rng.integers(0, 4)]]) for _ in range(n_jobs)]
    specs = [fac.job(c, n, k, variant=k % 2) for k, (c, n) in enumerate(plan)]
    # plain: every (circuit, size) alone on an idle prover
    plain = {}
    with frames.FramePool(1) as pool:
        for c in CIRCUITS:
            for sz in sizes:
                n = FRAME_SIZES[sz]
                sp = fac.job(c, n, 99999)
                pool.result(pool.submit(*sp))
                lat = []
                for _ in range(3):
                    r = pool.result(pool.submit(*sp))
                    lat.append(r.t_end - r.t_submit)
                plain[(c, n)] = sorted(lat)[1]
    plain_lat = np.array([plain[p] for p in plan])
    # solo proofs of a sample of the jobs: each served proof must equal them
    sample = list(range(0, n_jobs, max(1, n_jobs // 20)))
    with frames.FramePool(1) as pool:
        solo = {k: pool.result(pool.submit(*specs[k])).proof for k in sample}
    out = {"workload": "%d jobs, Poisson %.0f jobs/s, circuits uniform over %s, sizes uniform over "
                       "%s macroblocks" % (n_jobs, rate, "/".join(CIRCUITS),
                                           "/".join(str(FRAME_SIZES[s]) for s in sizes)),
           "offered_jobs_per_s": rate,
           "plain_p50_ms": round(float(np.percentile(plain_lat, 50)) * 1e3, 3),
           "plain_p99_ms": round(float(np.percentile(plain_lat, 99)) * 1e3, 3),
           "by_config": {}}
    for label, kw in variants:
        with frames.FramePool(**kw) as pool:
            for k in range(16):                                # warm every prover's buffers
                pool.result(pool.submit(*fac.job("gray", FRAME_SIZES["1080p"], 90000 + k)))
            t0 = time.perf_counter()
            p0 = pool.now()
            jobs = []
            for k, sp in enumerate(specs):
                while time.perf_counter() - t0 < arrive[k]:
                    pass
                jobs.append(pool.submit(*sp))
            res = [pool.result(j) for j in jobs]
            mem = pool.memory() if kw.get("batched") else None
        lat = np.array([r.t_end - r.t_submit for r in res])
        span = max(r.t_end for r in res) - p0
        rec = {"frames_per_s": round(n_jobs / span, 2),
               "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
               "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
               "mean_batch": round(float(np.mean([r.batch for r in res])), 2),
               "proofs_equal_solo": "%d/%d" % (sum(res[k].proof == solo[k] for k in sample),
                                               len(sample))}
        if mem is not None:
            rec["peak_admitted_GB"] = round(mem["peak"] / 1e9, 2)
            rec["budget_GB"] = round(mem["budget"] / 1e9, 1)
        out["by_config"][label] = rec
        print("bench: served %s p50 %.2f p99 %.2f ms" % (label, rec["p50_ms"], rec["p99_ms"]),
              file=sys.stderr, flush=True)
    best = min(out["by_config"], key=lambda k: out["by_config"][k]["p99_ms"])
    out["best_config"] = best
    for k in ("frames_per_s", "p50_ms", "p99_ms"):
        out[k] = out["by_config"][best][k]
    return out


def bench_served_sweep(rates=(200, 400, 800, 1600), n_jobs: int = 600, seed: int = 6,
                       variants=None):
    """C5 at rising Poisson rates: per pool and rate, the frames/s delivered
    and the latency percentiles, and the highest rate whose p99 stays within
    twice the plain p99 of the same job mix (each job alone on an idle
    prover)."""
    from truepix_amd import frames
    import numpy as np
    if variants is None:
        variants = [("stream-16", dict(n_workers=16)),
                    ("lockstep-2x64", dict(n_workers=2, batched=True, max_batch=64))]
    fac = _FrameFactory(seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = ["240p", "480p", "720p", "1080p"]
    plan = [(CIRCUITS[rng.integers(0, 4)], FRAME_SIZES[sizes[rng.integers(0, 4)]])
            for _ in range(n_jobs)]
    specs = [fac.job(c, n, k, variant=k % 2) for k, (c, n) in enumerate(plan)]
    plain = {}
    with frames.FramePool(1) as pool:
        for key in sorted(set(plan)):
            sp = fac.job(key[0], key[1], 99999)
            pool.result(pool.submit(*sp))
            lat = []
            for _ in range(3):
                r = pool.result(pool.submit(*sp))
                lat.append(r.t_end - r.t_submit)
            plain[key] = sorted(lat)[1]
    plain_p99 = float(np.percentile([plain[p] for p in plan], 99)) * 1e3
    out = {"workload": "%d jobs per rate, Poisson arrivals, circuits uniform over %s, sizes "
                       "uniform over %s macroblocks" % (n_jobs, "/".join(CIRCUITS),
                                                        "/".join(str(FRAME_SIZES[s]) for s in sizes)),
           "plain_p99_ms": round(plain_p99, 3), "by_config": {}}
    for label, kw in variants:
        rows = {}
        with frames.FramePool(**kw) as pool:
            for k in range(16):
                pool.result(pool.submit(*fac.job("gray", FRAME_SIZES["1080p"], 90000 + k)))
            for rate in rates:
                arrive = np.cumsum(np.random.Generator(np.random.PCG64(seed + rate))
                                   .exponential(1.0 / rate, size=n_jobs))
                t0 = time.perf_counter()
                p0 = pool.now()
                jobs = []
                for k, sp in enumerate(specs):
                    while time.perf_counter() - t0 < arrive[k]:
                        pass
                    jobs.append(pool.submit(*sp))
                res = [pool.result(j) for j in jobs]
                lat = np.array([r.t_end - r.t_submit for r in res]) * 1e3
                span = max(r.t_end for r in res) - p0
                rows[str(rate)] = {"frames_per_s": round(n_jobs / span, 1),
                                   "p50_ms": round(float(np.percentile(lat, 50)), 3),
                                   "p99_ms": round(float(np.percentile(lat, 99)), 3)}
                print("bench: sweep %s %d/s p99 %.1f ms" % (label, rate, rows[str(rate)]["p99_ms"]),
                      file=sys.stderr, flush=True)
        ok = [int(r) for r, v in rows.items() if v["p99_ms"] <= 2 * plain_p99]
        out["by_config"][label] = {
            "by_rate": rows,
            "max_rate_p99_within_2x_plain": max(ok) if ok else None,
            "frames_per_s_at_that_rate": rows[str(max(ok))]["frames_per_s"] if ok else None}
    return out


def bench_yuv(width: int = 1920, height: int = 1080, reps: int = 200, cpu: bool = True):
    """A 1080p frame's planes, resident on the device, packed into the
    prover's 8040 rows of 384 bytes and written back (tp_yuv_pack /
    tp_yuv_unpack), event-timed on the launch stream; bytes = read + write.
    CPU baseline: the oracle's numpy restatement of the reference's host
    packing (VideoEdition/Frame_Block_Extraction.py:67-94, 220-240)."""
    import numpy as np
    import torch
    from truepix_amd import yuv
    rng = np.random.default_rng(7)
    y = rng.integers(0, 256, (height, width), dtype=np.uint8)
    cb = rng.integers(0, 256, (height // 2, width // 2), dtype=np.uint8)
    cr = rng.integers(0, 256, (height // 2, width // 2), dtype=np.uint8)
    ty, tcb, tcr = (torch.from_numpy(a).cuda() for a in (y, cb, cr))
    n = yuv.copies(width, height)
    rows = torch.empty((n, 384), dtype=torch.uint8, device="cuda")
    out = {}
    for name, fn in (("pack", lambda: yuv.macroblocks_device(ty, tcb, tcr, out=rows)),
                     ("unpack", lambda: yuv.frame_device(rows, width, height))):
        for _ in range(5):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        nbytes = 2.0 * n * 384
        out[name] = {"us_per_frame": round(us, 2), "gbs": round(nbytes / us / 1e3, 1)}
    rec = {"frame": "%dx%d" % (width, height), "copies": n, "bytes_per_frame": 2 * n * 384, **out}
    if cpu:
        from oracle import yuv_oracle
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < 3.0:
            yuv_oracle.pack(y, cb, cr)
            k += 1
        dt = (time.perf_counter() - t0) / k
        rec["cpu_baseline"] = {"value": round(dt * 1e6, 1), "unit": "us per frame", "cores": 1,
                               "kind": "port", "sample": "%d packs of the same frame" % k}
    return rec


def bench_orion_commit(copies: int = 1350, reps: int = 5, seed: int = 777, cpu: bool = True):
    """The signer's Orion commitment of one SD frame (1350 macroblocks: n = 2^20
    coefficients per polynomial, 128 columns of 8192, codewords of 14,090):
    encode both polynomials, hash every row's 64 column pairs, build both
    Merkle trees.  CPU: the reference's own layout, mask, encode and Merkle
    code compiled from source (oracle/_ref), single-threaded, one commit."""
    import numpy as np
    from truepix_amd import orion
    x = np.random.Generator(np.random.PCG64(7)).integers(0, 256, size=(copies, 384), dtype=np.uint8)
    roots = orion.commit(x, seed)                       # builds the code's graphs once
    walls, devs = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        m, k, ms = orion.commit(x, seed, timing=True)
        walls.append(time.perf_counter() - t0)
        devs.append(ms)
    walls.sort()
    devs.sort()
    out = {"workload": "Orion commit of one frame: %d copies, n = %d per polynomial, masked + mask"
                       % (copies, (1 << max(0, (copies - 1).bit_length())) * 512),
           "ms_wall": round(walls[len(walls) // 2] * 1e3, 3),
           "ms_device": round(devs[len(devs) // 2], 3), "roots": list(roots)}
    if cpu:
        try:
            sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
            from oracle.oracle import RefOrion, Oracle
            if RefOrion.available():
                impl, kind, what = RefOrion(), "reference", "oracle/_ref (the reference's code)"
                t0 = time.perf_counter()
                rm, rk, _ = impl.commit(x, seed)
            else:
                impl, kind, what = Oracle(), "port", "oracle/fp2_oracle.c or_orion_commit"
                t0 = time.perf_counter()
                rm, rk, _ = impl.orion_commit(x, seed)
            dt = time.perf_counter() - t0
            out["cpu_baseline"] = {"ms": round(dt * 1e3, 1), "cores": 1, "kind": kind,
                                   "sample": "one commit at %d copies, %s" % (copies, what),
                                   "roots_equal": [rm, rk] == list(roots)}
            out["speedup_vs_cpu"] = round(dt * 1e3 / out["ms_wall"], 1)
        except Exception as exc:  # the baseline must not hide the device result
            out["cpu_baseline"] = {"error": str(exc)}
    out["open"] = bench_orion_open(x, seed, cpu=cpu)
    out["open_nizk"] = bench_orion_open_nizk(x, seed, cpu=cpu)
    return out


def bench_orion_open_nizk(x, seed: int, reps: int = 5, cpu: bool = True):
    """The NIZK opening of the frame's masked polynomial with the Virgo
    code-switching proof that closes it (linearPC.cpp:700-928): the opening
    as above with the transcript's 1222 positions, then 1222 more positions,
    the encoding circuit's layers proved on the device (two launches, every
    layer at once) and the verifier's claim chain.  CPU: the reference's own
    generate_circuit / zk_prover / zk_verifier checks for the same proof
    (oracle/_ref), or the oracle's restatement, one core."""
    import numpy as np
    from truepix_amd import orion
    copies = x.shape[0]
    pc = 1 << max(0, (copies - 1).bit_length())
    lg = (pc * 512).bit_length() - 1
    r = np.random.Generator(np.random.PCG64(11)).integers(0, (1 << 61) - 1, size=(lg, 2),
                                                         dtype=np.uint64)
    orion.opening_nizk_virgo(x, seed, r, True)
    walls, devs = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        proof, _, words = orion.opening_nizk_virgo(x, seed, r, True)
        walls.append(time.perf_counter() - t0)
        devs.append(orion.virgo_stats()[0])
    walls.sort()
    devs.sort()
    layers = orion.virgo_stats()[2]
    out = {"workload": "NIZK Orion opening + Virgo code-switching proof: %d copies, message %d, "
                       "%d sumcheck layers" % (copies, pc * 512 // 128, layers),
           "ms_wall": round(walls[len(walls) // 2] * 1e3, 3),
           "virgo_ms_device": round(devs[len(devs) // 2], 4),
           "virgo_verdict": int(words[-1]), "virgo_words": int(words.size)}
    if cpu:
        try:
            from oracle.oracle import RefVirgo, Oracle
            o = orion.parse_opening(proof)
            args = (o["com_root"], pc * 512, True, r, o["combined_root"], o["combined_message"])
            t0 = time.perf_counter()
            if RefVirgo.available():
                ref, kind = RefVirgo().transcript(*args), "reference"
            else:
                ref, kind = Oracle().orion_virgo(*args), "port"
            dt = time.perf_counter() - t0
            out["cpu_baseline"] = {"ms": round(dt * 1e3, 1), "cores": 1, "kind": kind,
                                   "sample": "the code-switching proof of the same opening",
                                   "transcript_equal": bool(np.array_equal(ref, words))}
            out["virgo_speedup_vs_cpu"] = round(dt * 1e3 / out["virgo_ms_device"], 1)
        except Exception as exc:
            out["cpu_baseline"] = {"error": str(exc)}
    return out


def bench_orion_open(x, seed: int, reps: int = 5, n_queries: int = 1222, cpu: bool = True):
    """An opening of the frame's masked polynomial (tensor_product_protocol
    up to its Virgo proof): re-encode and hash as the commit does, combine the
    128 codewords with r0, hash and tree the combined codeword, gather the
    1222 queried columns and both trees' siblings, serialize the
    OpeningProof; wall time including the proof's copy to the host.  CPU: the
    reference's own code for the same (oracle/_ref), one core."""
    import hashlib
    import numpy as np
    from truepix_amd import orion
    copies = x.shape[0]
    pc = 1 << max(0, (copies - 1).bit_length())
    lg = (pc * 512).bit_length() - 1
    rng = np.random.Generator(np.random.PCG64(11))
    r = rng.integers(0, (1 << 61) - 1, size=(lg, 2), dtype=np.uint64)
    proof, _ = orion.opening(x, seed, r, [0], use_group1=True)
    cw = orion.parse_opening(proof)["cw_len"]
    q = rng.integers(0, cw, size=n_queries, dtype=np.int64)
    walls = []
    for _ in range(reps):
        t0 = time.perf_counter()
        proof, ans = orion.opening(x, seed, r, q, use_group1=True)
        walls.append(time.perf_counter() - t0)
    walls.sort()
    out = {"workload": "Orion opening of one frame's masked polynomial: %d copies, %d queries"
                       % (copies, n_queries),
           "ms_wall": round(walls[len(walls) // 2] * 1e3, 3), "proof_bytes": len(proof),
           "proof_sha256": hashlib.sha256(proof).hexdigest()}
    if cpu:
        try:
            from oracle.oracle import RefOrion, Oracle
            t0 = time.perf_counter()
            if RefOrion.available():
                ref, kind = RefOrion().open(x, seed, r, q, True), "reference"
            else:
                ref, kind = Oracle().orion_open(x, seed, r, q, True)[0], "port"
            dt = time.perf_counter() - t0
            out["cpu_baseline"] = {"ms": round(dt * 1e3, 1), "cores": 1, "kind": kind,
                                   "sample": "one opening, the same frame, point and queries",
                                   "proof_equal": ref == proof}
            out["speedup_vs_cpu"] = round(dt * 1e3 / out["ms_wall"], 1)
        except Exception as exc:
            out["cpu_baseline"] = {"error": str(exc)}
    return out


def bench_fft_gkr(sizes=(7, 17, 20), reps: int = 3, cpu_lg: int = 17, cpu: bool = True):
    """fft_gkr (OZ/src/VPD/fft_circuit_GKR.cpp): circuit build + every layer
    protocol + the verifier's checks, per lg_size (2^lg coefficients; the
    TruePix pipeline runs lg 7).  Wall time per run and the device span
    (HIP events from the first to the last kernel).  CPU: the oracle's
    restatement, single-threaded, at cpu_lg."""
    from truepix_amd import fft_gkr
    out = {"workload": "fft_gkr: IFFT + 64-point evaluation circuit over 2^lg Fp2 coefficients, "
                       "addition / mult / intermediate / lg IFFT layer sumchecks, reference "
                       "rand() challenges, verifier checks", "by_lg": {}}
    for lg in sizes:
        fft_gkr.run(lg)
        walls, devs = [], []
        ok = True
        for _ in range(reps):
            t0 = time.perf_counter()
            okk, v_time, psize, p_time = fft_gkr.run(lg)
            walls.append(time.perf_counter() - t0)
            devs.append(p_time)
            ok = ok and okk
        walls.sort()
        devs.sort()
        out["by_lg"][str(lg)] = {"ms_wall": round(walls[len(walls) // 2] * 1e3, 3),
                                 "ms_device": round(devs[len(devs) // 2] * 1e3, 3),
                                 "verified": ok, "proof_size": psize}
    if cpu:
        from oracle.oracle import Oracle
        t0 = time.perf_counter()
        r = Oracle().fft_gkr(cpu_lg, with_polys=False)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"lg": cpu_lg, "ms": round(dt * 1e3, 1), "verified": bool(r["ok"]),
                               "cores": 1, "kind": "port",
                               "sample": "one full run at lg %d, oracle/fp2_oracle.c or_fft_gkr" % cpu_lg}
        if str(cpu_lg) in out["by_lg"]:
            out["speedup_vs_cpu_at_lg%d" % cpu_lg] = round(dt * 1e3 / out["by_lg"][str(cpu_lg)]["ms_wall"], 1)
    return out


def cpu_baseline(nv_sample: int = 24, nthreads: int = 1, min_seconds: float = 10.0):
    from oracle.oracle import Oracle
    from truepix_amd.sumcheck import synthetic_tables
    from truepix_amd.fiatshamir import FiatShamir
    from truepix_amd.extfield import _raw
    o = Oracle()
    f, g = synthetic_tables(nv_sample, seed=7)
    fs = FiatShamir()
    fs.put(["c2", 7])

    def chal(k, p):
        fs.put([_raw(int(p[0]), int(p[1])), _raw(int(p[2]), int(p[3])), _raw(int(p[4]), int(p[5]))])
        r = fs.rand_ext("c2-%d" % k)
        return r.real, r.imag
    reps = 0
    t0 = time.perf_counter()
    while True:
        o.sc2_prove(f, g, chal, nthreads=nthreads)
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= min_seconds:
            break
    return {"value": round(reps * alg_bytes(nv_sample) / dt / 1e9, 3), "unit": "GB/s",
            "cores": nthreads, "kind": "port",
            "sample": "%d full %d-round proves over 2 x 2^%d Fp2 tables, %.1f s on %d thread(s) "
                      "(oracle/fp2_oracle.c, -O3, OpenMP)" % (reps, nv_sample, nv_sample, dt,
                                                              nthreads)}


def hbm_ceiling(reps: int = 5):
    """HBM rates plain kernels reach on this GPU, beside the published peak:
    a 256 MiB -> 256 MiB copy (read + write) and a sum over 512 MiB of int64
    (read), median of `reps` launches timed with events."""
    import torch
    n = 1 << 26                                    # 2^26 int64 = 512 MiB
    x = torch.ones(n, dtype=torch.int64, device="cuda")
    y = torch.empty(n // 2, dtype=torch.int64, device="cuda")

    def timed(fn):
        ts = []
        for _ in range(reps + 1):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        return sorted(ts[1:])[reps // 2]
    t_copy = timed(lambda: y.copy_(x[: n // 2]))
    t_read = timed(lambda: x.sum())
    out = {"copy_gbs": round(2 * (n // 2) * 8 / t_copy / 1e9, 1),
           "read_gbs": round(n * 8 / t_read / 1e9, 1),
           "how": "torch copy_ of 256 MiB (read + write) and sum of 512 MiB int64, median of %d" % reps}
    del x, y
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run the two kernels whose HBM counters the traffic figure "
                         "compares: a 256 MiB elementwise copy (x + 0, read 256 MiB, write "
                         "256 MiB) and one 24-round prove")
    ap.add_argument("--batched-configs", default="",
                    help="comma list of batched pool labels to run (default: all), e.g. "
                         "lockstep-1x256")
    ap.add_argument("--skip", default="", help="comma list of side measurements to skip "
                    "(gkr_frame, gkr_frame_python_driver, batched, served, served_sweep, "
                    "fft_gkr, orion, yuv, rs_encode)")
    a = ap.parse_args(argv)
    a.skip = set(x for x in a.skip.split(",") if x)
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the proofs here are single-GPU work "
              "(frames and tables shard by frame, with no cross-GPU exchange); "
              "--gpus %d is not supported" % a.gpus, file=sys.stderr)
        return 2

    import torch
    from truepix_amd import native_fp2_bridge as nb
    from truepix_amd import sumcheck as sc
    nb.require()

    f, g = sc.synthetic_tables(NV, seed=2024)
    df, dg = sc.to_device(f), sc.to_device(g)
    del f, g
    stream = torch.cuda.Stream().cuda_stream            # the kernels' stream (not the null stream)
    if a.traffic_probe:
        x = torch.zeros(1 << 25, dtype=torch.int64, device="cuda")
        y = torch.add(x, 0)                              # calibration: 256 MiB in, 256 MiB out
        torch.cuda.synchronize()
        sc.prove_deg2_device(df.data_ptr(), dg.data_ptr(), NV, seed=2024, stream=stream)
        torch.cuda.synchronize()
        print(json.dumps({"traffic_probe": True, "copy_bytes": int(y.numel() * 8)}))
        return 0
    for _ in range(a.warmup):
        sc.prove_deg2_device(df.data_ptr(), dg.data_ptr(), NV, seed=2024, stream=stream)

    # timed region: exactly K steps, synchronized on both sides
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        proof = sc.prove_deg2_device(df.data_ptr(), dg.data_ptr(), NV, seed=2024, stream=stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt / a.steps * 1e3
    value = alg_bytes(NV) / (dt / a.steps) / 1e9

    # the same proves with every launch down to the mid launch queued ahead
    # of its transcript step (untimed, outside the timed region above), for
    # comparison with the default of launching after each step
    sc.set_ahead(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        sc.prove_deg2_device(df.data_ptr(), dg.data_ptr(), NV, seed=2024, stream=stream)
    torch.cuda.synchronize()
    ms_queued = (time.perf_counter() - t0) / a.steps * 1e3
    sc.set_ahead(False)

    # roofline per launch class: events around each launch, in separate
    # (untimed) runs so event recording does not perturb `value`
    sc.set_timing(True)
    reps = max(3, a.steps // 4)
    acc = {c: {"ms": 0.0, "bytes": 0.0, "launches": 0} for c in sc.CLASSES}
    for _ in range(reps):
        sc.prove_deg2_device(df.data_ptr(), dg.data_ptr(), NV, seed=2024, stream=stream)
        for c, v in sc.class_stats().items():
            for key in v:
                acc[c][key] += v[key]
    tail_prof = sc.tail_profile()
    mid_prof = sc.mid_profile()
    sc.set_timing(False)
    classes = {}
    for c, v in acc.items():
        if v["launches"]:
            classes[c] = {"kernel": SC2_KERNELS[c], "launches_per_prove": v["launches"] // reps,
                          "avg_launch_us": round(v["ms"] / v["launches"] * 1e3, 2),
                          "bytes_per_launch": round(v["bytes"] / v["launches"], 1),
                          "ms_per_prove": round(v["ms"] / reps, 4),
                          "achieved_gbs": round(v["bytes"] / (v["ms"] / 1e3) / 1e9, 1)}
    if "tail" in classes:
        classes["tail"]["per_round"] = tail_prof
    if "mid" in classes:
        classes["mid"]["per_round"] = mid_prof
    # the roofline is the HBM one: the class moving the most bytes per prove
    dom = max((c for c in classes if c not in ("tail", "mid")), key=lambda c: acc[c]["bytes"])
    achieved = acc[dom]["bytes"] / (acc[dom]["ms"] / 1e3) / 1e9

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: uniform canonical Fp2 tables from numpy PCG64(2024)",
        "config": {"workload": "degree-2 sumcheck, 2 x 2^24 Fp2 tables, 24 rounds, "
                               "TruePix transcript challenges (fft_gkr sumcheck_phase1_update)",
                   "table_entries": 1 << NV, "rounds": NV, "proves_per_s": round(1e3 / ms_per_step, 2),
                   "ms_per_step_queued_launches": round(ms_queued, 4),
                   "alg_bytes_per_prove": alg_bytes(NV)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": measured_traffic(SC2_KERNELS[dom]),
                     "valu": measured_valu(SC2_KERNELS[dom]),
                     "kernel": SC2_KERNELS[dom],
                     "avg_launch_us": classes[dom]["avg_launch_us"],
                     "launches_per_prove": classes[dom]["launches_per_prove"],
                     "alg_bytes_per_launch": classes[dom]["bytes_per_launch"],
                     "classes": classes,
                     "ceiling": hbm_ceiling(),
                     "note": "first: the 256 block products of the input on the int8 matrix "
                             "cores (one read of the input) and their reduction, from which the "
                             "host takes rounds 0 to 3; fold01: one pass folding the input by "
                             "r_0 .. r_3 into T_4 with round 4's sums; pair: passes folding by "
                             "two challenges (inputs outside the block pass); fold: streaming "
                             "fold+round launches; mid: 128 workgroups, each folding its own "
                             "slice of the tables in LDS, workgroup 0 gathering their sums for "
                             "the host, which answers the challenges through mapped memory; "
                             "tail: one workgroup proving the last rounds, the host answering "
                             "through mapped memory (latency-bound); mid and tail take two "
                             "rounds per host exchange (per_round figures are per exchange); "
                             "ceiling: what plain torch kernels reach on this GPU now"},
        "check": {"round0_poly": [int(x) for x in proof.polys[0][:2]]},
    }
    def batched():
        if not a.batched_configs:
            return bench_batched()
        want = a.batched_configs.split(",")
        return bench_batched(variants=[c for c in BATCHED_CONFIGS if c[0] in want])

    for key, fn in (("gkr_frame", bench_frames_plain), ("gkr_frame_python_driver", bench_gkr_frame),
                    ("batched", batched), ("served", bench_served),
                    ("served_sweep", bench_served_sweep)):
        if key in a.skip:
            continue
        try:
            line[key] = fn()
        except Exception as exc:
            line[key] = {"error": "%s: %s" % (type(exc).__name__, exc)}
        print("bench: %s done" % key, file=sys.stderr, flush=True)
        if key == "gkr_frame" and not a.no_cpu_baseline and "error" not in line[key]:
            try:
                cb = cpu_frame_baseline()
                cb["gpu_speedup"] = round(line[key]["frames_per_s"] / cb["value"], 1)
                line[key]["cpu_baseline"] = cb
            except Exception as exc:
                line[key]["cpu_baseline"] = {"error": "%s: %s" % (type(exc).__name__, exc)}
    if "fft_gkr" not in a.skip:
        try:
            line["fft_gkr"] = bench_fft_gkr(cpu=not a.no_cpu_baseline)
        except Exception as exc:
            line["fft_gkr"] = {"error": "%s: %s" % (type(exc).__name__, exc)}
        print("bench: fft_gkr done", file=sys.stderr, flush=True)
    if "orion" not in a.skip:
        try:
            line["orion_commit"] = bench_orion_commit(cpu=not a.no_cpu_baseline)
        except Exception as exc:
            line["orion_commit"] = {"error": "%s: %s" % (type(exc).__name__, exc)}
        print("bench: orion_commit done", file=sys.stderr, flush=True)
    if "yuv" not in a.skip:
        try:
            line["yuv_layout"] = bench_yuv(cpu=not a.no_cpu_baseline)
        except Exception as exc:
            line["yuv_layout"] = {"error": "%s: %s" % (type(exc).__name__, exc)}
        print("bench: yuv_layout done", file=sys.stderr, flush=True)
    if "rs_encode" not in a.skip:
        try:
            line["rs_encode"] = bench_rs_encode(cpu=not a.no_cpu_baseline)
        except Exception as exc:
            line["rs_encode"] = {"error": str(exc)}
    if not a.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline()
            n = host_threads()
            line["cpu_baseline"]["all_cores"] = cpu_baseline(nthreads=n, min_seconds=4.0)
        except Exception as exc:  # the baseline must not hide the GPU result
            line["cpu_baseline"] = {"error": str(exc)}
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
