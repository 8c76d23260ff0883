"""The drop-ins for the reference's Orion commit and prove programs
(truepix_amd/orion_cli.py, for linearPC_multi_commit_zk and
linearPC_multi_prove_zk) run as the TruePix scripts run them: same
arguments, environment and files.  The manifest's roots and both NIZK proof
files equal the oracle's (pinned to the reference) for the same frame, seed
and point; the printed lines parse as the scripts parse them; a wrong
manifest root makes the prove step fail as the reference's does."""
import os
import subprocess
import sys

import numpy as np
import pytest

from truepix_amd.extfield import Fp2
from truepix_amd.fp2_link import write_point_file

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = (1 << 61) - 1


def _run(tmp_path, env, *args):
    print("orion_cli", args, flush=True)
    e = dict(os.environ)
    e.update(env)
    e["PYTHONPATH"] = ROOT + os.pathsep + e.get("PYTHONPATH", "")
    return subprocess.run([sys.executable, "-m", "truepix_amd.orion_cli"] + list(args), cwd=tmp_path,
                          env=e, capture_output=True, text=True, timeout=300)


def _root_bytes(hex_root):
    words = [int(hex_root[16 * i:16 * i + 16], 16) for i in range(4)]
    out = bytearray(32)
    for w, off in zip(words, (8, 0, 24, 16)):
        out[off:off + 8] = w.to_bytes(8, "little")
    return bytes(out)


def _parse_times(stdout):
    """TP/libTruePix/fp2_link.py:119-141 and run_truepix_signer.py:52-58."""
    total, commit = 0.0, None
    for line in stdout.splitlines():
        if "Commit time(mask)" in line:
            commit = float(line.split()[-1])
        if "Open time" in line or "Commit time" in line:
            total += float(line.split()[-1])
    return commit, total


@pytest.mark.parametrize("copies,layout", [(16, (384, 384, 128)), (40, (385, -1, 0))])
def test_commit_and_prove_programs(bridge, oracle, tmp_path, copies, layout):
    cs, ci, cv = layout
    x = np.random.Generator(np.random.PCG64(copies)).integers(0, 256, size=(copies, cs), dtype=np.uint8)
    with open(tmp_path / "input.txt", "w") as fh:
        for row in x:
            fh.write(" ".join(str(int(v)) for v in row) + "\n")
    env = {"TRUEPIX_PROTOCOL": "fp2-linked-v1", "TRUEPIX_COPY_SIZE": str(cs),
           "TRUEPIX_CONST_IDX": str(ci), "TRUEPIX_CONST_VAL": str(cv)}
    c = _run(tmp_path, env, "commit", str(copies), "0")
    assert c.returncode == 0, c.stderr
    commit_t, _ = _parse_times(c.stdout)
    assert commit_t is not None and commit_t > 0
    import json
    man = json.load(open(tmp_path / "orion_commit.json"))
    seed = int(open(tmp_path / "orion_mask_seed.txt").read().split()[0])
    oracle.orion_set_layout(cs, ci, cv)
    try:
        m, k, cw = oracle.orion_commit(x, seed)
        assert (man["masked_root"], man["mask_root"]) == (m, k)
        assert (man["copy_size"], man["const_idx"], man["const_val"], man["zk"]) == (cs, ci, cv, True)
        lg = man["lg_n"]
        pt = np.random.Generator(np.random.PCG64(copies + 1)).integers(0, P, size=(lg, 2), dtype=np.uint64)
        write_point_file(str(tmp_path / "gkr_point.json"), [Fp2(int(a), int(b)) for a, b in pt])
        env["TRUEPIX_GKR_POINT_FILE"] = str(tmp_path / "gkr_point.json")
        p = _run(tmp_path, env, "prove", str(copies), "0")
        assert p.returncode == 0, p.stdout + p.stderr
        assert "check pass" in p.stdout and "mle_consistent=1" in p.stdout
        _, total = _parse_times(p.stdout)
        assert total > 0
        sizes = [int(l.split()[-2]) for l in p.stdout.splitlines() if "Proof size for tensor IOP" in l]
        assert len(sizes) == 2
        r_orion = np.ascontiguousarray(pt[::-1])
        for group, name, root in ((True, "orion_proof_masked.bin", m), (False, "orion_proof_mask.bin", k)):
            q = oracle.orion_draw_queries(_root_bytes(root), man["n"], group, cw, r_orion)
            want, _ = oracle.orion_open(x, seed, r_orion, np.asarray(q, np.int64), use_group1=group)
            assert open(tmp_path / name, "rb").read() == want, name
        # a manifest that names another root: the prove step fails
        man["mask_root"] = "0" * 64
        json.dump(man, open(tmp_path / "orion_commit.json", "w"))
        bad = _run(tmp_path, env, "prove", str(copies), "0")
        assert bad.returncode == 1 and "check notpass" in bad.stdout and "root_mask=0" in bad.stdout
    finally:
        oracle.orion_set_layout(384, 384, 128)
