"""Whole-frame proving on the device: one call per TruePix frame, and a pool
of provers for batched and served operation.

A frame is one run of ``CircuitProverNIZK.run``
(TP/libTruePix/circuitnizkvec_vole.py:265-501) for a one-layer circuit, as
``run_truepix_prover.run_truepix`` proves an edited frame
(TP/run_truepix_prover.py:218-423).  The whole protocol (transcript, VOLE
commitments, interpolation, proof bytes) runs in the library next to the
kernels (``tp_pool_*`` in include/truepix_hip.h); this module packs the
circuit, inputs, VOLE correlations and commitment binding for it.

``FramePool(n)`` keeps n provers (host threads with their own HIP streams)
taking frames from one FIFO queue: ``submit`` as frames arrive, ``wait`` /
``result`` per frame.  ``FramePool(n, batched=True)`` keeps n lockstep batch
drivers instead: each takes every queued frame that fits its batch size and
the pool's device-memory budget and proves them together, one launch per
protocol step across the batch (frames arriving meanwhile form the next
batch).  The proof of a frame does not depend on which worker proves it, on
the batch it is in or on what else is in flight (each frame has its own
transcript).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from truepix_amd import native_fp2_bridge as nb
from truepix_amd.extfield import BASE_PRIME_61, _raw
from truepix_amd.fp2_link import commit_binding, read_commit_manifest
from truepix_amd.vole import ROLE_PROVER, read_vole_file

P = BASE_PRIME_61
_U64P = ctypes.POINTER(ctypes.c_uint64)
_I32P = ctypes.POINTER(ctypes.c_int)


def _lib():
    nb.require()
    return nb.LIB


def _p(a: np.ndarray, typ):
    return a.ctypes.data_as(typ)


class FrameCircuit:
    """A one-layer circuit resident on the device (tp_circuit_create), from a
    pypws-shaped description (TP/libTruePix/parse_pws.py:8-64)."""

    def __init__(self, parsed, muxbits: Optional[Sequence[int]] = None):
        from truepix_amd.gkr import parse_layers, refuse_mux
        input_layer, layers = parse_layers(parsed)
        if len(layers) != 1:
            raise ValueError("the frame prover takes one-layer circuits (got %d layers)" % len(layers))
        refuse_mux(layers)
        lay = layers[0]
        muxlen = (int(lay.mux.max()) if lay.n_gates else 0) + 1
        bits = list(muxbits) if muxbits is not None else [0] * muxlen
        self.muxbits = bits
        self.layer = lay
        self.input_layer = list(input_layer)
        self.n_inputs = len(input_layer)
        self.n_private = sum(1 for c in input_layer if c is None)
        self.in_bits = lay.in_bits
        self.out_bits = lay.out_bits
        slot = np.full(self.n_inputs, -1, np.int32)
        const = np.zeros(self.n_inputs, np.uint64)
        k = 0
        for w, c in enumerate(input_layer):
            if c is None:
                slot[w] = k
                k += 1
            else:
                const[w] = int(c) % P
        mux = np.ascontiguousarray([bits[m] if m < len(bits) else 0 for m in lay.mux], np.int32)
        type_idx = np.ascontiguousarray(lay.type_idx(bits), np.int32)
        lib = _lib()
        self._lib = lib
        self.handle = lib.tp_circuit_create(lay.n_gates, _p(lay.op_idx, _I32P), _p(mux, _I32P),
                                            _p(type_idx, _I32P), _p(lay.in0, _I32P),
                                            _p(lay.in1, _I32P), self.n_inputs, _p(slot, _I32P),
                                            _p(const, _U64P))
        if not self.handle:
            raise RuntimeError("tp_circuit_create failed: %s" % nb.last_error())

    def close(self) -> None:
        if getattr(self, "handle", None):
            self._lib.tp_circuit_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def vole_prover_arrays(path: str):
    """(U.re, U.im, V.re, V.im) of a prover EVOL file with one repetition
    (TP/libTruePix/ExtVOLEGen.py:156-187)."""
    blob = read_vole_file(path)
    if blob["role"] != ROLE_PROVER:
        raise ValueError("Not an ExtVOLE prover file for truepix")
    if blob["security_repetitions"] != 1:
        raise ValueError("the frame prover takes one VOLE security repetition")
    (ur, ui), (vr, vi) = blob["streams"]
    return tuple(np.ascontiguousarray(a, np.uint64) for a in (ur, ui, vr, vi))


@dataclass
class FrameResult:
    proof: bytes
    input_point: list           # Fp2 per variable: nInBits line point, then nCopyBits
    input_value: object         # Fp2, the input MLE at input_point
    t_submit: float
    t_start: float
    t_end: float
    phases_ms: dict = field(default_factory=dict)
    outputs: Optional[np.ndarray] = None    # [copies_pad][2^out_bits] when kept
    batch: int = 1                          # frames in the launch set that proved it
    footprint: int = 0                      # device table bytes admission counted

    @property
    def latency(self) -> float:
        return self.t_end - self.t_submit


PHASES = ("eval_download", "absorb_outputs", "tables", "early_rounds", "late_rounds", "tail",
          "input_mle", "proof_bytes")


class FramePool:
    """n_workers device provers fed from one queue (tp_pool_*).

    batched=True: n_workers lockstep batch drivers (tp_pool_create_batched),
    each proving up to max_batch frames per launch set, admitted in arrival
    order while their table bytes fit budget_bytes (0: three quarters of the
    free device memory), with the frames' transcript work on host_threads
    threads."""

    def __init__(self, n_workers: int = 1, batched: bool = False, max_batch: int = 256,
                 budget_bytes: int = 0, host_threads: int = 16):
        self._lib = _lib()
        if batched:
            self.handle = self._lib.tp_pool_create_batched(int(n_workers), int(max_batch),
                                                           int(budget_bytes), int(host_threads))
        else:
            self.handle = self._lib.tp_pool_create(int(n_workers))
        if not self.handle:
            raise RuntimeError("tp_pool_create failed: %s" % nb.last_error())
        self.n_workers = int(n_workers)
        self.batched = bool(batched)
        # job id -> its FrameCircuit, so a caller's circuit object outlives its
        # jobs even when the caller drops it (the library also holds a
        # reference on the device circuit per job)
        self._circuits = {}

    BATCH_PROFILE = ("batches", "frames", "stage_wall_ms", "launch_wall_ms", "head_wall_ms",
                     "head_wait_sum_ms", "head_cpu_sum_ms", "prefix_cpu_sum_ms", "rounds_wall_ms",
                     "step_wait_ms", "step_host_wall_ms", "rest_wall_ms", "claim_device_ms",
                     "fold_device_ms")

    def tuning(self) -> dict:
        """The lockstep schedule tuner: phase (2 = settled), the knobs in use
        and the measured cost of every candidate."""
        buf = (ctypes.c_double * 9)()
        if self._lib.tp_pool_tuning(self.handle, buf, 9) < 0:
            raise RuntimeError("tp_pool_tuning failed")
        return {"phase": int(buf[0]), "claim_items": int(buf[1]), "fold_cap": int(buf[2]),
                "claim_ns_per_term": [round(buf[3 + i], 4) for i in range(3)],
                "fold_ns_per_byte": [round(buf[6 + i], 6) for i in range(3)]}

    def fix_schedule(self, claim_items: int, fold_cap: int) -> None:
        """Fix the lockstep schedule (the tuner stops); proofs do not change."""
        if self._lib.tp_pool_fix_schedule(self.handle, int(claim_items), int(fold_cap)) != 0:
            raise RuntimeError("tp_pool_fix_schedule failed: %s" % nb.last_error())

    def batch_profile(self) -> dict:
        """Where the batches of a batched pool spent their time (accumulated)."""
        buf = (ctypes.c_double * 32)()
        n = self._lib.tp_pool_batch_profile(self.handle, buf, 32)
        if n < 0:
            raise RuntimeError("tp_pool_batch_profile failed")
        return {k: buf[i] for i, k in enumerate(self.BATCH_PROFILE[:n])}

    def memory(self) -> dict:
        """Admission state: bytes admitted now, budget, peak; batches and frames."""
        v = [ctypes.c_uint64() for _ in range(5)]
        if self._lib.tp_pool_memory(self.handle, *[ctypes.byref(x) for x in v]) != 0:
            raise RuntimeError("tp_pool_memory failed")
        return dict(zip(("in_use", "budget", "peak", "batches", "frames"), (x.value for x in v)))

    def now(self) -> float:
        return self._lib.tp_pool_now(self.handle)

    def submit(self, circuit: FrameCircuit, inputs, pid: int, vole, binding: list,
               keep_outputs: bool = False) -> int:
        """Queue one frame.  inputs: [n_copies][n_private] ints (uint8 values
        are sent as bytes); vole: (U.re, U.im, V.re, V.im) arrays; binding: the
        commit_binding list (str / int items)."""
        job, _keep = self._job(circuit, inputs, pid, vole, binding, keep_outputs)
        jid = self._lib.tp_pool_submit(self.handle, ctypes.byref(job))
        if jid <= 0:
            raise RuntimeError("tp_pool_submit failed: %s" % nb.last_error())
        self._circuits[int(jid)] = circuit
        return int(jid)

    def submit_yuv(self, circuit: FrameCircuit, y, cb, cr, pid: int, vole, binding: list,
                   extra=None, keep_outputs: bool = False) -> int:
        """Queue one frame given as YUV 4:2:0 planes (uint8 numpy arrays, or
        torch tensors already on the GPU): its per-copy rows are built on the
        device (``yuv.macroblocks_device``; ``extra`` appends a byte per copy,
        e.g. ``yuv.crop_bits`` for the crop circuit) and handed to the prover
        there, so a frame crosses PCIe once, as planes."""
        import torch
        from truepix_amd import yuv

        def dev(a):
            if isinstance(a, torch.Tensor):
                return a if a.is_cuda else a.cuda()
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint8)).cuda()

        ex = None if extra is None else dev(np.asarray(extra, dtype=np.uint8).reshape(-1)
                                          if not isinstance(extra, torch.Tensor) else extra)
        rows = yuv.macroblocks_device(dev(y), dev(cb), dev(cr), ex)
        torch.cuda.current_stream().synchronize()   # the rows are complete before submit
        return self.submit(circuit, rows, pid, vole, binding, keep_outputs)

    def submit_many(self, specs) -> List[int]:
        """Queue frames at once, in order: specs are submit()'s argument
        tuples.  A batch driver admits them together."""
        built = [self._job(*sp) for sp in specs]
        arr = (nb.FrameJob * max(1, len(built)))(*[b[0] for b in built])
        ids = (ctypes.c_int64 * max(1, len(built)))()
        if self._lib.tp_pool_submit_many(self.handle, arr, len(built), ids) != 0:
            raise RuntimeError("tp_pool_submit_many failed: %s" % nb.last_error())
        for k, sp in enumerate(specs):
            self._circuits[int(ids[k])] = sp[0]
        return [int(ids[k]) for k in range(len(built))]

    def _job(self, circuit: FrameCircuit, inputs, pid: int, vole, binding: list,
             keep_outputs: bool = False):
        """The C job description and the arrays it points into."""
        if not getattr(circuit, "handle", None):
            raise ValueError("the circuit is closed")
        on_device = type(inputs).__module__.startswith("torch") and getattr(inputs, "is_cuda", False)
        if on_device:
            # rows resident on the GPU (uint8, contiguous, complete): the
            # library copies them device to device at submit
            import torch
            if inputs.dtype != torch.uint8 or inputs.dim() != 2 or \
                    inputs.shape[1] != circuit.n_private or not inputs.is_contiguous():
                raise ValueError("device inputs must be a contiguous uint8 [n_copies][%d] tensor"
                                 % circuit.n_private)
            data, nbytes, n_copies, ptr = inputs, 1, int(inputs.shape[0]), inputs.data_ptr()
        else:
            arr = np.asarray(inputs)
            if arr.ndim != 2 or arr.shape[1] != circuit.n_private:
                raise ValueError("inputs must be [n_copies][%d]" % circuit.n_private)
            if arr.dtype == np.uint8 or (arr.size and arr.min() >= 0 and arr.max() <= 255):
                data = np.ascontiguousarray(arr, dtype=np.uint8)
                nbytes = 1
            else:
                data = np.ascontiguousarray(np.asarray(arr, dtype=object) % P, dtype=np.uint64)
                nbytes = 8
            n_copies, ptr = arr.shape[0], data.ctypes.data
        ur, ui, vr, vi = (np.ascontiguousarray(a, np.uint64) for a in vole)
        kinds = np.array([1 if isinstance(b, str) else 0 for b in binding], np.int32)
        ints = np.array([0 if isinstance(b, str) else int(b) for b in binding], np.uint64)
        strs = (ctypes.c_char_p * max(1, len(binding)))(
            *[b.encode("utf-8") if isinstance(b, str) else None for b in binding])
        job = nb.FrameJob(circuit.handle, n_copies, circuit.n_private, nbytes, ptr, int(pid),
                          ur.size, _p(ur, _U64P), _p(ui, _U64P), _p(vr, _U64P), _p(vi, _U64P),
                          len(binding), _p(kinds, _I32P), _p(ints, _U64P), strs,
                          1 if keep_outputs else 0, 1 if on_device else 0)
        return job, (data, ur, ui, vr, vi, kinds, ints, strs)

    def wait(self, job: int, timeout: Optional[float] = None) -> bool:
        rc = self._lib.tp_pool_wait(self.handle, job, -1.0 if timeout is None else float(timeout))
        if rc < 0:
            raise RuntimeError("frame %d failed: %s" % (job, self._lib.tp_pool_error(self.handle,
                                                                                  job).decode()))
        return rc == 1

    def result(self, job: int, release: bool = True) -> FrameResult:
        self.wait(job)
        lib, h = self._lib, self.handle
        n = ctypes.c_size_t()
        ptr = ctypes.c_void_p()
        if lib.tp_pool_proof_view(h, job, ctypes.byref(ptr), ctypes.byref(n)) != 0:
            raise RuntimeError("tp_pool_proof_view failed")
        proof = ctypes.string_at(ptr.value, n.value) if n.value else b""
        pt = np.zeros(2 * 64, np.uint64)
        npt = ctypes.c_int()
        val = np.zeros(2, np.uint64)
        times = (ctypes.c_double * 3)()
        ph = (ctypes.c_double * 8)()
        if lib.tp_pool_info(h, job, _p(pt, _U64P), 64, ctypes.byref(npt), _p(val, _U64P), times,
                            ph) != 0:
            raise RuntimeError("tp_pool_info failed")
        bsz = ctypes.c_int()
        fp = ctypes.c_uint64()
        lib.tp_pool_job_stats(h, job, ctypes.byref(bsz), ctypes.byref(fp))
        outs = None
        cnt = ctypes.c_size_t()
        if lib.tp_pool_outputs(h, job, None, 0, ctypes.byref(cnt)) == 0 and cnt.value:
            outs = np.zeros(cnt.value, np.uint64)
            lib.tp_pool_outputs(h, job, _p(outs, _U64P), cnt.value, ctypes.byref(cnt))
        res = FrameResult(
            proof=proof,
            input_point=[_raw(int(pt[2 * k]), int(pt[2 * k + 1])) for k in range(npt.value)],
            input_value=_raw(int(val[0]), int(val[1])),
            t_submit=times[0], t_start=times[1], t_end=times[2],
            phases_ms={k: ph[i] for i, k in enumerate(PHASES)},
            outputs=outs, batch=bsz.value, footprint=fp.value)
        if release:
            self.release(job)
        return res

    def release(self, job: int) -> None:
        """Forget a finished job (its proof, outputs and circuit reference)."""
        self._lib.tp_pool_release(self.handle, job)
        self._circuits.pop(int(job), None)

    def close(self) -> None:
        if getattr(self, "handle", None):
            self._lib.tp_pool_destroy(self.handle)
            self.handle = None
            self._circuits.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FrameProverNIZK:
    """CircuitProverNIZK's interface (TP/libTruePix/circuitnizkvec_vole.py:21-501)
    for one-layer circuits, proving the whole frame inside the library.

    Reads vole_ext_prover_<pid>.json from ``vole_dir`` and the commitment
    manifest from ``manifest_path`` (orion_commit.json), like the reference."""

    def __init__(self, parsed, nCopies: int, pool: Optional[FramePool] = None):
        self.parsed = parsed
        self.nCopies = int(nCopies)
        self.circuit = FrameCircuit(parsed)
        self.pool = pool
        self.pid = None
        self.vole_dir = "."
        self.manifest_path = None
        self.final_input_point = None
        self.final_input_value = None
        self.ckt_outputs = None
        self.timings = None

    @classmethod
    def from_pws(cls, parsed, nCopies):
        return parsed[0], cls(parsed, nCopies)

    def run(self, inputs, muxbits=None) -> bytes:
        if muxbits is not None:
            raise ValueError("the frame prover takes circuits without mux bits")
        pid = os.getpid() if self.pid is None else int(self.pid)
        rows = np.asarray(inputs)
        if rows.shape[1] == self.circuit.n_inputs and self.circuit.n_private != self.circuit.n_inputs:
            priv_cols = [w for w, c in enumerate(self.circuit.input_layer) if c is None]
            rows = rows[:, priv_cols]
        vole = vole_prover_arrays(os.path.join(self.vole_dir, "vole_ext_prover_%d.json" % pid))
        binding = commit_binding(read_commit_manifest(self.manifest_path))
        own = self.pool is None
        pool = FramePool(1) if own else self.pool
        try:
            jid = pool.submit(self.circuit, rows, pid, vole, binding, keep_outputs=True)
            res = pool.result(jid)
        finally:
            if own:
                pool.close()
        out_w = 1 << self.circuit.out_bits
        self.ckt_outputs = res.outputs.reshape(-1, out_w)
        self.final_input_point = res.input_point
        self.final_input_value = res.input_value
        self.timings = dict(res.phases_ms)
        return res.proof
