"""GPU parity of fft_gkr (A16 circuit build + A18 layer protocols, OZ/src/VPD/
fft_circuit_GKR.cpp): every round polynomial, the claim after every layer, the
final alpha / beta, the circuit layers, the pass flag and the proof size equal
the oracle's (oracle/fp2_oracle.c or_fft_gkr) and the reference's golden
values (tests/golden/fft_gkr_ref.json); at the benchmark size, the
reference's own verifier checks (restated on the host) accept the device run."""
import hashlib
import json
import os

import numpy as np
import pytest

from truepix_amd import fft_gkr

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fft_gkr_ref.json")


def gold_cases():
    with open(GOLD) as fh:
        return json.load(fh)["cases"]


def layer_digests(layers, lg):
    n = 1 << lg
    raw = layers.tobytes()
    return {"ifft_out": hashlib.sha256(raw[:16 * n]).hexdigest(),
            "scaled": hashlib.sha256(raw[16 * n:32 * n]).hexdigest(),
            "sums": hashlib.sha256(raw[32 * n:32 * n + 64 * 16]).hexdigest()}


@pytest.mark.parametrize("case", gold_cases(), ids=lambda c: "lg%d" % c["lg_size"])
def test_fft_gkr_matches_reference_and_oracle(bridge, oracle, case):
    lg = case["lg_size"]
    print("fft_gkr lg", lg, flush=True)
    d = fft_gkr.trace(lg, layers=True)
    assert d["ok"] == case["ok"] == 1
    assert d["proof_size"] == case["proof_size"]
    assert [int(x) for x in d["claims"]] == case["claims"]
    assert layer_digests(d["layers"], lg) == case["layer_sha256"]
    o = oracle.fft_gkr(lg)
    assert d["n_polys"] == o["n_polys"]
    assert np.array_equal(d["polys"], o["polys"])


def test_fft_gkr_larger_sizes_vs_oracle(bridge, oracle):
    for lg in (14, 16):
        print("fft_gkr lg", lg, flush=True)
        d = fft_gkr.trace(lg, layers=True)
        o = oracle.fft_gkr(lg)
        assert d["ok"] == o["ok"] == 1
        assert np.array_equal(d["claims"], o["claims"])
        assert np.array_equal(d["polys"], o["polys"])
        assert np.array_equal(d["layers"], o["layers"])


def test_fft_gkr_benchmark_size_verifies(bridge):
    lg = 20
    ok, v_time, proof_size, p_time = fft_gkr.run(lg)
    assert ok
    rounds = (lg + 6) + lg + 2 * lg * lg
    assert proof_size == 48 * rounds + 48 * lg * (lg + 1) // 2
    assert p_time > 0 and v_time >= 0


def test_fft_gkr_cli_writes_reference_log_line(bridge, tmp_path):
    out = tmp_path / "log_fftgkr.txt"
    assert fft_gkr.main(["7", str(out)]) == 0
    v, size, p = out.read_text().split()
    assert int(size) == 7008          # the reference's proof size at lg 7
    float(v), float(p)


def test_graph_replays_equal_eager_runs(bridge):
    """The captured launch sequence, replayed several times with eager runs of
    other sizes in between, gives the eager result every time."""
    try:
        fft_gkr.set_graph(False)
        eager = fft_gkr.trace(9, layers=True)
        fft_gkr.set_graph(True)
        for rep in range(3):
            got = fft_gkr.trace(9, layers=True)
            for key in ("polys", "claims", "layers"):
                assert np.array_equal(got[key], eager[key]), (rep, key)
            assert got["ok"] == 1
            fft_gkr.set_graph(False)
            other = fft_gkr.trace(5 + rep)
            assert other["ok"] == 1
            fft_gkr.set_graph(True)
    finally:
        fft_gkr.set_graph(True)
