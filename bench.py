#!/usr/bin/env python3
"""Benchmark: degree-2 sumcheck prover over two resident 2^24-entry Fp2 tables
(24 rounds, TruePix transcript challenges) on one MI355X.

One step = one full 24-round prove (round sums, fused folds, host transcript)
from tables already resident in HBM.  Prints ONE JSON line:
  value      whole-job fold HBM throughput: algorithmic bytes of all rounds
             (32 n for round 0, then 48 x the input length of every fused
             fold+round launch, plus the final fold) / wall time per step;
  roofline   the fused fold+round kernel: its algorithmic bytes / its summed
             HIP-event time (events on the stream it runs on), against 8 TB/s;
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


def alg_bytes(nv: int) -> float:
    """Algorithmic HBM bytes of one nv-round prove (both tables, 16 B per Fp2)."""
    n = 1 << nv
    total = 32.0 * n                       # round 0 reads both tables
    m = n
    for _ in range(1, nv):                 # fused launches read T_{k-1} (m) and write m/2
        total += 48.0 * m
        m //= 2
    total += 32.0 * 2                      # final fold reads the last pair
    return total


def bench_rs_encode(steps: int = 5, cols: int = 64, lc: int = 20, lo: int = 22):
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
    bfly = cols * lc * (1 << (lo - 1))
    del dc, do
    return {"workload": "RS encode 64 x 2^20 -> 2^22 Fp2 (rate 1/4)", "ms": round(ms, 3),
            "alg_GBps": round(alg / ms / 1e6, 1), "Gbutterflies_per_s": round(bfly / ms / 1e6, 2),
            "alg_bytes": alg}


def bench_gkr_frame(frames: int = 5, circuit: str = "gray", n_copies: int = 1350):
    """Plain GKR prove of one frame (C1 shape: gray, 1350 macroblocks, 384
    samples each) with this package's prover; proof bytes as the reference's."""
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
            "sample": "%d full %d-round proves over 2 x 2^%d Fp2 tables, %.1f s single-threaded "
                      "(oracle/fp2_oracle.c, -O3)" % (reps, nv_sample, nv_sample, dt)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args(argv)
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

    # roofline of the fused fold+round kernel: events around each launch,
    # in separate (untimed) runs so event recording does not perturb `value`
    sc.set_timing(True)
    k_ms = k_bytes = 0.0
    k_n = 0
    for _ in range(max(3, a.steps // 4)):
        sc.prove_deg2_device(df.data_ptr(), dg.data_ptr(), NV, seed=2024, stream=stream)
        ms, by, n = sc.stats()
        k_ms += ms
        k_bytes += by
        k_n += n
    sc.set_timing(False)
    achieved = k_bytes / (k_ms / 1e3) / 1e9

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
                   "alg_bytes_per_prove": alg_bytes(NV)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": None,
                     "kernel": "k_sc2_fold_round",
                     "avg_launch_us": round(k_ms / k_n * 1e3, 2),
                     "alg_bytes_per_launch": round(k_bytes / k_n, 1)},
        "check": {"round0_poly": [int(x) for x in proof.polys[0][:2]]},
    }
    try:
        line["gkr_frame"] = bench_gkr_frame()
    except Exception as exc:
        line["gkr_frame"] = {"error": str(exc)}
    try:
        line["rs_encode"] = bench_rs_encode()
    except Exception as exc:
        line["rs_encode"] = {"error": str(exc)}
    if not a.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline()
        except Exception as exc:  # the baseline must not hide the GPU result
            line["cpu_baseline"] = {"error": str(exc)}
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
