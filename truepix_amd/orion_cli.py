"""Drop-ins for the reference's Orion commit and prove programs, running on
the device.

TruePix runs the ZK Orion programs as separate processes
(``TP/libTruePix/fp2_link.py:639-652``, ``orion_binary``):

  - ``linearPC_multi_commit_zk <copies> 0`` from the signer
    (``TP/run_truepix_signer.py:28-60``;
    ``OZ/examples/multivariate_PC_test_commit.cpp:265-331``): packs input.txt,
    draws a mask seed into ``orion_mask_seed.txt``, commits the masked and
    mask polynomials and writes ``orion_commit.json``;
  - ``linearPC_multi_prove_zk <copies> 0`` from the prover
    (``TP/run_truepix_prover.py:360-405``;
    ``OZ/examples/multivariate_PC_test_prove.cpp:298-390``): recommits,
    checks both roots against the manifest, opens both polynomials at the
    GKR input point (NIZK, each opening closed by its Virgo code-switching
    proof) into ``orion_proof_masked.bin`` and ``orion_proof_mask.bin``, and
    checks that the two values differ by the plaintext MLE at that point.

``python -m truepix_amd.orion_cli commit <copies> 0`` and ``... prove
<copies> 0`` take the same arguments, environment (``TRUEPIX_PROTOCOL``,
``TRUEPIX_COPY_SIZE`` / ``CONST_IDX`` / ``CONST_VAL``,
``TRUEPIX_ORION_COMMIT_FILE``, ``TRUEPIX_MASK_SEED_FILE``,
``TRUEPIX_GKR_POINT_FILE``) and files, write the same manifest and the same
proof bytes, print the lines the TruePix scripts parse (``Commit time(mask)``,
``Open time(...)``, ``Proof size for tensor IOP``, ``check pass``) and exit 0
or 1 as the programs do.  The commit and both openings run in
``libtruepix_hip.so``.

The figure after ``Proof size for tensor IOP`` is this opening's size as
written (the OpeningProof file) plus the code-switching proof's round
polynomials (48 bytes each); the reference's figure is an interactive-proof
accounting that also counts its FRI-based VPD, which is not built here.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

PROTOCOL_ID = "fp2-linked-v1"
PROTOCOL_ALIASES = ("fp2-linked-v1", "truepix")
LAYOUT_ID = "copy-major-384-const128-pad512-v1"
PADDED_COPY_SIZE = 512
P = (1 << 61) - 1


def _err(msg: str) -> int:
    print(msg, file=sys.stderr)
    return 1


def _env(name: str, default: str) -> str:
    v = os.environ.get(name)
    return v if v else default


def _enabled() -> bool:
    return os.environ.get("TRUEPIX_PROTOCOL") in PROTOCOL_ALIASES


def _args(argv):
    """(copies, userandom) as truepix_gkr.h:67-84 parses them, or None."""
    if len(argv) < 2:
        _err("usage: <copies> <userandom>")
        return None
    try:
        copies = int(argv[0])
        if copies <= 0:
            raise ValueError
    except ValueError:
        _err("invalid copies '%s': expected a positive integer" % argv[0])
        return None
    if argv[1] not in ("0", "1"):
        _err("invalid userandom '%s': expected 0 or 1" % argv[1])
        return None
    return copies, int(argv[1])


def make_layout(copies: int):
    """The packing and sizes of make_layout (truepix_gkr.h:107-150), or None."""
    cs = int(os.environ.get("TRUEPIX_COPY_SIZE", 384))
    ci = int(os.environ.get("TRUEPIX_CONST_IDX", 384))
    cv = int(os.environ.get("TRUEPIX_CONST_VAL", 128))
    if cs <= 0 or cs > PADDED_COPY_SIZE:
        _err("TruePix: invalid TRUEPIX_COPY_SIZE=%d (padded copy is %d)" % (cs, PADDED_COPY_SIZE))
        return None
    if ci >= PADDED_COPY_SIZE:
        _err("TruePix: TRUEPIX_CONST_IDX=%d out of range" % ci)
        return None
    if 0 <= ci < cs:
        _err("TruePix: TRUEPIX_CONST_IDX=%d overlaps the %d values read from input.txt; GKR and "
             "Orion would pack different polynomials" % (ci, cs))
        return None
    padded = 1 << max(0, (copies - 1).bit_length())
    n = padded * PADDED_COPY_SIZE
    return {"copies": copies, "padded_copies": padded, "copy_size": cs,
            "padded_copy_size": PADDED_COPY_SIZE, "const_idx": ci, "const_val": cv, "n": n,
            "lg_n": n.bit_length() - 1}


def load_samples(path: str, lay):
    """[copies][copy_size] sample values from input.txt (load_coefs,
    truepix_gkr.h:156-205: the first copy_size values of each of the first
    copies lines), or None."""
    rows = []
    try:
        with open(path) as fh:
            for line in fh:
                if len(rows) >= lay["copies"]:
                    break
                vals = [int(v) for v in line.split()[:lay["copy_size"]]]
                if any(v < 0 for v in vals):
                    _err("TruePix: negative value on line %d of %s" % (len(rows) + 1, path))
                    return None
                if len(vals) != lay["copy_size"]:
                    _err("TruePix: copy %d has %d values, expected %d" % (len(rows), len(vals),
                                                                        lay["copy_size"]))
                    return None
                rows.append(vals)
    except OSError:
        _err("TruePix: cannot open %s for reading" % path)
        return None
    if len(rows) < lay["copies"]:
        _err("TruePix: %s has %d lines, expected %d" % (path, len(rows), lay["copies"]))
        return None
    arr = np.asarray(rows, dtype=np.int64)
    if arr.size and arr.max() > 255:
        _err("TruePix: sample values above 255 in %s" % path)
        return None
    print("TruePix layout: C=%d padded_copies=%d N=%d lg_N=%d copy_size=%d const[%d]=%d"
          % (lay["copies"], lay["padded_copies"], lay["n"], lay["lg_n"], lay["copy_size"],
             lay["const_idx"], lay["const_val"]))
    return arr.astype(np.uint8)


def _packing(lay):
    return (lay["copy_size"], lay["const_idx"], lay["const_val"])


def write_manifest(lay, masked_root: str, mask_root: str, zk: bool = True) -> str:
    """orion_commit.json in write_manifest's text form (truepix_gkr.h:428-455)."""
    path = _env("TRUEPIX_ORION_COMMIT_FILE", "orion_commit.json")
    proto = os.environ.get("TRUEPIX_PROTOCOL")
    proto = proto if proto in PROTOCOL_ALIASES else PROTOCOL_ID
    body = ("{\n"
            '  "protocol": "%s",\n  "layout": "%s",\n  "zk": %s,\n  "copies": %d,\n'
            '  "padded_copies": %d,\n  "copy_size": %d,\n  "padded_copy_size": %d,\n'
            '  "const_idx": %d,\n  "const_val": %d,\n  "n": %d,\n  "lg_n": %d,\n'
            '  "masked_root": "%s",\n  "mask_root": "%s"\n}\n'
            % (proto, LAYOUT_ID, "true" if zk else "false", lay["copies"], lay["padded_copies"],
               lay["copy_size"], lay["padded_copy_size"], lay["const_idx"], lay["const_val"],
               lay["n"], lay["lg_n"], masked_root, mask_root))
    with open(path, "w") as fh:
        fh.write(body)
    print("TruePix: wrote commit manifest %s (zk=%s masked_root=%s mask_root=%s)"
          % (path, "true" if zk else "false", masked_root, mask_root))
    return path


def read_point(lg_n: int):
    """The GKR input point of TRUEPIX_GKR_POINT_FILE's plaintext companion
    (load_gkr_point, truepix_gkr.h:244-290): lg_n (re, im) pairs, or None."""
    env = os.environ.get("TRUEPIX_GKR_POINT_FILE")
    if env is None:
        _err("TruePix: TRUEPIX_GKR_POINT_FILE is not set")
        return None
    if env.endswith(".json"):
        path = env[:-5] + ".txt"
    elif env.endswith(".txt"):
        path = env
    else:
        path = env + ".txt"
    try:
        with open(path) as fh:
            tok = fh.read().split()
    except OSError:
        _err("TruePix: cannot open GKR point file %s" % path)
        return None
    if not tok or int(tok[0]) != lg_n:
        _err("TruePix: GKR point has %s coordinates, expected %d" % (tok[0] if tok else 0, lg_n))
        return None
    vals = [int(t) for t in tok[1:1 + 2 * lg_n]]
    if len(vals) != 2 * lg_n:
        _err("TruePix: GKR point file %s truncated" % path)
        return None
    if any(v >= P for v in vals):
        _err("TruePix: GKR point coordinate is not reduced")
        return None
    print("TruePix: loaded GKR point from %s" % path)
    return np.asarray(vals, dtype=np.uint64).reshape(lg_n, 2)


def commit_main(argv) -> int:
    """linearPC_multi_commit_zk (multivariate_PC_test_commit.cpp:265-331)."""
    from truepix_amd import orion
    a = _args(argv)
    if a is None:
        return 1
    copies, userandom = a
    if not _enabled():
        return _err("this drop-in runs the TruePix protocol only (set TRUEPIX_PROTOCOL)")
    if userandom:
        return _err("fp2-linked-v1 requires shared input.txt (userandom=0)")
    lay = make_layout(copies)
    if lay is None:
        return 1
    x = load_samples("input.txt", lay)
    if x is None:
        return 1
    seed = int.from_bytes(os.urandom(8), "little")
    seed_path = _env("TRUEPIX_MASK_SEED_FILE", "orion_mask_seed.txt")
    with open(seed_path, "w") as fh:
        fh.write("%d\n" % seed)
    print("Running fused mask commit...")
    t0 = time.perf_counter()
    masked_root, mask_root = orion.commit(x, seed, layout=_packing(lay))
    print("Commit time(mask): %f" % (time.perf_counter() - t0))
    write_manifest(lay, masked_root, mask_root, True)
    return 0


def _mle(x, lay, r):
    """The plaintext MLE of the packed coefficients at the GKR point (mle_gkr,
    truepix_gkr.h:212-229, r[b] on bit b of the index), on the device."""
    from truepix_amd import native_fp2_bridge as nb
    from truepix_amd.extfield import Fp2
    coefs = np.zeros(lay["n"], dtype=np.uint64)
    rows = coefs.reshape(lay["padded_copies"], PADDED_COPY_SIZE)
    rows[:lay["copies"], :lay["copy_size"]] = x
    if lay["const_idx"] >= 0:
        rows[:lay["copies"], lay["const_idx"]] = lay["const_val"]
    v = nb.mle_eval_base(coefs, [Fp2(int(re), int(im)) for re, im in r])
    return int(v.real), int(v.imag)


def prove_main(argv) -> int:
    """linearPC_multi_prove_zk (multivariate_PC_test_prove.cpp:298-390)."""
    from truepix_amd import orion
    a = _args(argv)
    if a is None:
        return 1
    copies, userandom = a
    if not _enabled():
        return _err("this drop-in runs the TruePix protocol only (set TRUEPIX_PROTOCOL)")
    if userandom:
        return _err("fp2-linked-v1 requires shared input.txt (userandom=0)")
    lay = make_layout(copies)
    if lay is None:
        return 1
    man_path = _env("TRUEPIX_ORION_COMMIT_FILE", "orion_commit.json")
    try:
        with open(man_path) as fh:
            man = json.load(fh)
    except (OSError, ValueError):
        return _err("TruePix: cannot open commit manifest %s; run the commit phase "
                    "(run_fennel_signer.py) first" % man_path)
    if not man.get("zk"):
        return _err("TruePix: ZK Orion cannot open a non-ZK commitment")
    if any(man.get(k) != lay[k] for k in lay):
        return _err("TruePix: layout does not match the commit manifest")
    x = load_samples("input.txt", lay)
    if x is None:
        return 1
    seed_path = _env("TRUEPIX_MASK_SEED_FILE", "orion_mask_seed.txt")
    try:
        with open(seed_path) as fh:
            seed = int(fh.read().split()[0])
    except (OSError, ValueError, IndexError):
        return _err("TruePix: cannot open mask seed file %s; run the commit phase "
                    "(run_fennel_signer.py) first" % seed_path)
    r = read_point(lay["lg_n"])
    if r is None:
        return 1
    r_orion = np.ascontiguousarray(r[::-1])   # reverse_point: the opening is MSB-first
    pack = _packing(lay)
    print("Running fused mask commit...")
    t0 = time.perf_counter()
    roots = orion.commit(x, seed, layout=pack)
    print("Commit time(mask): %f" % (time.perf_counter() - t0))
    mle_clear = _mle(x, lay, r)
    reports = []
    for group, label, expect, path in ((True, "masked", man.get("masked_root"), "orion_proof_masked.bin"),
                                       (False, "mask", man.get("mask_root"), "orion_proof_mask.bin")):
        root = roots[0] if group else roots[1]
        root_ok = root == expect
        if not root_ok:
            print("TruePix: %s Merkle root mismatch\n  committed: %s\n  manifest:  %s"
                  % (label, root, expect), file=sys.stderr)
        t0 = time.perf_counter()
        try:
            proof, value, words = orion.opening_nizk_virgo(x, seed, r_orion, group, layout=pack)
        except RuntimeError as exc:
            print(str(exc), file=sys.stderr)
            proof, value, words = None, (0, 0), np.zeros(0, np.uint64)
        dt = time.perf_counter() - t0
        ok = proof is not None
        if ok:
            with open(path, "wb") as fh:
                fh.write(proof)
            print("ORION_NIZK_PROOF_BYTES %d" % len(proof))
            print("ORION_NIZK_PROOF_FILE %s" % path)
            rounds = 0
            if words.size:
                bits = [int(b) for b in words[1:1 + int(words[0])]]
                rounds = 2 * sum(bits[:-1])
            print("Proof size for tensor IOP %d bytes" % (len(proof) + 48 * rounds))
        print("Open time(%s): %f" % (label, dt))
        print("Result: %s" % ("succ" if ok else "fail"))
        reports.append((ok, root_ok, value))
    (om, rm, vm), (ok2, rk, vk) = reports
    value = ((vm[0] - vk[0]) % P, (vm[1] - vk[1]) % P)
    consistent = value == mle_clear
    ok = om and ok2 and rm and rk and consistent
    print("TruePix prove: open_masked=%d open_mask=%d root_masked=%d root_mask=%d "
          "mle_consistent=%d -> ok=%d" % (om, ok2, rm, rk, consistent, ok))
    print("check pass" if ok else "check notpass")
    sys.stdout.flush()
    return 0 if ok else 1


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv or argv[0] not in ("commit", "prove"):
        return _err("usage: python -m truepix_amd.orion_cli {commit|prove} <copies> <userandom>")
    return (commit_main if argv[0] == "commit" else prove_main)(argv[1:])


if __name__ == "__main__":
    sys.exit(main())
