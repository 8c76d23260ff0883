"""CPU checks of the host protocol pieces against the golden fixtures written
from the reference (tests/golden/make_golden_*.py)."""
import bz2
import hashlib
import json
import os
import pickle

import numpy as np
import pytest

from truepix_amd import circuits
from truepix_amd.extfield import Fp2, proof_pickle_names
from truepix_amd.fiatshamir import FiatShamir, encode_items
from truepix_amd.gkr import interpolate_cubic, interpolate_quadratic, lagrange_coeffs, horner_eval
from truepix_amd.vole import ExtVOLECommit, ExtVOLEGenerator, read_vole_file

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = (1 << 61) - 1


def load_index():
    with open(os.path.join(GOLD, "golden_proofs.json")) as fh:
        return json.load(fh)


def golden_proof(entry):
    with open(os.path.join(GOLD, entry["proof_file"]), "rb") as fh:
        return bz2.decompress(fh.read())


@pytest.mark.parametrize("name", ["gray", "invert", "crop", "mask"])
def test_restated_circuits_match_pws(name):
    with open(os.path.join(GOLD, "circuits.json")) as fh:
        want = json.load(fh)[name]
    parsed = circuits.circuit(name)
    assert hashlib.sha256(repr(parsed).encode()).hexdigest() == want["sha256"]
    assert len(parsed[0]) == want["n_inputs"] and len(parsed[1]) == want["n_gates"]


# The packing TP/libTruePix/fp2_link.py:540-558 assigns to each PWS file
# (literal constants of the reference source): copy_size, const_idx, const_val.
REFERENCE_LAYOUTS = {"gray.pws": (384, 384, 128), "inv.pws": (384, 384, 255),
                     "mask.pws": (384, -1, 0), "Crop.pws": (385, -1, 0)}


@pytest.mark.parametrize("name,pws", [("gray", "gray.pws"), ("invert", "inv.pws"),
                                      ("mask", "mask.pws"), ("crop", "Crop.pws")])
def test_circuit_input_layouts_match_the_reference_packing(tmp_path, name, pws):
    """The restated circuits put their private inputs and their constant wire
    where the reference's commitment packing expects them, and an input.txt
    line of a frame has copy_size values."""
    from truepix_amd.fp2_link import infer_gkr_input_layout, input_layout, random_private_inputs
    parsed = circuits.circuit(name)
    want = REFERENCE_LAYOUTS[pws]
    assert input_layout(parsed) == want
    assert infer_gkr_input_layout("TP/pws/" + pws, str(tmp_path / "none.txt")) == want
    priv = random_private_inputs(name, parsed, 3, seed=1)
    with open(tmp_path / "input.txt", "w") as fh:
        for row in priv.tolist():
            fh.write(" ".join(map(str, row)) + "\n")
    assert infer_gkr_input_layout(pws, str(tmp_path / "input.txt")) == want
    # the constant sits right after the private samples, as the manifest says
    if want[1] >= 0:
        assert all(c is None for c in parsed[0][:want[1]])


def test_pws_parser_subset():
    text = "P V0 = I0 E\nP V1 = I1 E\nP V2 = 7 E\nP V3 = 9 E\nP O4 = V0 * V2 E\nP O5 = V1 minus V0 E\n"
    parsed = circuits.parse_pws_text(text)
    assert parsed == [[None, None, 7], [("MUL", 0, 2, None), ("SUB", 1, 0, None)]]
    with pytest.raises(ValueError):
        circuits.parse_pws_text("P V0 = I0 E\nP O1 = V0 / V0 E\n")


def test_vole_files_reproducible_and_parse(tmp_path):
    idx = load_index()
    pid = idx["pid"]
    gen = ExtVOLEGenerator(seed=idx["vole_seed"])
    gen.generate_deltas()
    gen.generate_vole_instances(500)
    pf, vf = tmp_path / "p.bin", tmp_path / "v.bin"
    gen.save_to_files(str(pf), str(vf))
    for mine, gold in ((pf, "vole_ext_prover_%d.json" % pid), (vf, "vole_ext_verifier_%d.json" % pid)):
        assert open(mine, "rb").read() == open(os.path.join(GOLD, gold), "rb").read()
    blob = read_vole_file(os.path.join(GOLD, "vole_ext_verifier_%d.json" % pid))
    assert blob["role"] == 2 and blob["n"] == 500
    # P = U Delta + V for every correlation
    pb = read_vole_file(os.path.join(GOLD, "vole_ext_prover_%d.json" % pid))
    (ur, ui), (vr, vi) = pb["streams"]
    (pr, pi), (dr, di) = blob["streams"][0], blob["deltas"][0]
    for k in range(0, 500, 37):
        u = Fp2(int(ur[k]), int(ui[k]))
        want = u * Fp2(dr, di) + Fp2(int(vr[k]), int(vi[k]))
        assert want == Fp2(int(pr[k]), int(pi[k]))


def test_vole_commit_roundtrip():
    pid = load_index()["pid"]
    pc, vc = ExtVOLECommit(), ExtVOLECommit()
    pc.load_from_file_prover(os.path.join(GOLD, "vole_ext_prover_%d.json" % pid))
    vc.load_from_file_verifier(os.path.join(GOLD, "vole_ext_verifier_%d.json" % pid))
    m1, m2 = Fp2(123, 456), Fp2(7, 9)
    t1, d1 = pc.commit(m1)
    t2, d2 = pc.commit(m2)
    p1, p2 = vc.change_commit(d1), vc.change_commit(d2)
    c = Fp2(31337, 4)
    z1, z2 = pc.pok_finish(m1, t1, c)
    assert vc.pok_check(z1, z2, c, p1)
    assert vc.est_eq_check(p1, p2, c, pc.est_eq(t1, t2, c)) is False or m1 == m2
    prod = m1 * m2
    t3, d3 = pc.commit(prod)
    p3 = vc.change_commit(d3)
    assert vc.prod_check((p1, p2, p3), pc.prod_finish((m1, m2, prod), (t1, t2, t3)))
    assert vc.est_val_check(p1, c, m1, pc.est_val(t1, c))
    assert not vc.est_val_check(p1, c, m2, pc.est_val(t1, c))


def test_proof_pickle_roundtrip_is_bytewise():
    """Unpickling a reference proof with the restricted loader and pickling it
    again with this package's Fp2 gives the same bytes."""
    for entry in load_index()["cases"]:
        if "proof_file" not in entry:
            continue
        raw = golden_proof(entry)
        (q, io, trans, ndb, rvs, rve, wdiv) = FiatShamir.unpack_proof(raw)
        with proof_pickle_names():
            again = pickle.dumps((q, list(io), list(trans), ndb, rvs, rve, wdiv, 4), -1)
        assert again == raw, entry["name"]


def test_proof_loader_rejects_foreign_globals():
    evil = pickle.dumps((P, [], [os.getcwd], None, None, None, None, 4), -1)
    with pytest.raises(pickle.UnpicklingError):
        FiatShamir.unpack_proof(evil)


def test_transcript_encoding_forms():
    assert encode_items([5, 6]) == b"\x01" + (2).to_bytes(4, "little") + (5).to_bytes(8, "little") \
        + (6).to_bytes(8, "little")
    assert encode_items([None]) == b"\x05"
    assert encode_items(["ab", 3]) == b"\x04" + (2).to_bytes(4, "little") + b"ab" + b"\x02" \
        + (3).to_bytes(8, "little")
    assert encode_items([Fp2(1, 2)]) == b"\x03" + (1).to_bytes(8, "little") + (2).to_bytes(8, "little")
    assert encode_items([-1]) == b"\x02" + ((1 << 64) - 1).to_bytes(8, "little")
    fs = FiatShamir()
    a = fs.rand_ext("x")
    assert 0 <= a.real < P and 0 <= a.imag < P and fs.rnum == 2


def test_interpolation_closed_forms():
    rng = np.random.default_rng(5)
    for _ in range(20):
        c = [Fp2(int(a), int(b)) for a, b in rng.integers(0, P, size=(4, 2))]
        f = lambda x: horner_eval(c, x)  # noqa: E731
        assert interpolate_cubic([f(0), f(1), f(P - 1), f(2)]) == c
        q = c[:3]
        g = lambda x: horner_eval(q, x)  # noqa: E731
        assert interpolate_quadratic([g(0), g(1), g(P - 1)]) == q
    for n in (2, 5, 10):
        c = [Fp2(int(a), int(b)) for a, b in rng.integers(0, P, size=(n, 2))]
        ys = [horner_eval(c, t) for t in range(n)]
        assert lagrange_coeffs(ys) == c
