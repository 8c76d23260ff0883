"""Pins the CPU oracle (oracle/fp2_oracle.c) before it is trusted as the checker.

Each function is checked against closed forms / brute force (always) and
against the reference's own sources compiled in place (oracle/_ref, when
/root/reference is present to build them).
"""
import ctypes

import numpy as np
import pytest

from oracle.oracle import RefNative, RefRS, RefSumcheck, _p, rand_fp2
from tests import fieldpy as F

P = F.P
EDGE = [0, 1, 2, P - 1, P - 2, (1 << 60), (1 << 32) - 1, 1 << 32, 1537228672809129301]


def out2():
    return (ctypes.c_uint64 * 2)()


def test_fp2_mul_closed_form(oracle):
    rng = np.random.default_rng(11)
    cases = [((a, b), (c, d)) for a in EDGE for b in EDGE[:4] for c in EDGE[:5] for d in EDGE[3:7]]
    cases += [(x, y) for x, y in zip(F.rand_elems(rng, 300), F.rand_elems(rng, 300))]
    for x, y in cases:
        o = out2()
        oracle.fp2_mul(x[0], x[1], y[0], y[1], o)
        assert (o[0], o[1]) == F.fmul(x, y)


def test_fp2_inverse_and_constants(oracle):
    rng = np.random.default_rng(12)
    for x in F.rand_elems(rng, 50) + [(1, 0), (0, 1), (P - 1, 5)]:
        o = out2()
        oracle.fp2_inv(x[0], x[1], o)
        assert F.fmul(x, (o[0], o[1])) == (1, 0)
    # Defs.half / Defs.third (TP/libTruePix/defs.py:18-19) are inverses of 2 and 3
    assert 2 * 1152921504606846976 % P == 1
    assert 3 * 1537228672809129301 % P == 1


@pytest.mark.parametrize("tp", list(range(13)))
def test_gate_truth_tables(oracle, tp):
    for x in (0, 1):
        for y in (0, 1):
            o = out2()
            oracle.gatefn(tp, x, 0, y, 0, o)
            assert (o[0], o[1]) == (F.GATE_TRUTH[tp](x, y) % P, 0), (tp, x, y)
    # multilinear: affine in x for fixed y and in y for fixed x
    rng = np.random.default_rng(100 + tp)
    x0, x1, y, t = F.rand_elems(rng, 4)
    def g(a, b):
        o = out2()
        oracle.gatefn(tp, a[0], a[1], b[0], b[1], o)
        return (o[0], o[1])
    xt = F.fadd(x0, F.fmul(t, F.fsub(x1, x0)))
    want = F.fadd(g(x0, y), F.fmul(t, F.fsub(g(x1, y), g(x0, y))))
    assert g(xt, y) == want


def test_fold_is_mle(oracle):
    rng = np.random.default_rng(13)
    for nv in (1, 2, 4, 6):
        vals = F.rand_elems(rng, 1 << nv)
        z = F.rand_elems(rng, nv)
        re, im = F.to_soa(vals)
        for j in range(nv):
            n_out = re.size // 2
            orr = np.zeros(n_out, np.uint64)
            oi = np.zeros(n_out, np.uint64)
            oracle.fold(_p(re), _p(im), n_out, z[j][0], z[j][1], _p(orr), _p(oi))
            re, im = orr, oi
        assert (int(re[0]), int(im[0])) == F.mle(vals, z)


def test_compute_beta_closed_form(oracle):
    rng = np.random.default_rng(14)
    for nz in (1, 3, 7):
        z = F.rand_elems(rng, nz)
        init = F.rand_elems(rng, 1)[0]
        zr, zi = F.to_soa(z)
        orr = np.zeros(1 << nz, np.uint64)
        oi = np.zeros(1 << nz, np.uint64)
        oracle.compute_beta(_p(zr), _p(zi), nz, init[0], init[1], _p(orr), _p(oi))
        for x in range(1 << nz):
            assert (int(orr[x]), int(oi[x])) == F.fmul(init, F.chi(x, z))


def test_mle_eval_base_bruteforce(oracle):
    rng = np.random.default_rng(15)
    for nz in (1, 2, 5, 8):
        ins = rng.integers(0, 1 << 63, size=1 << nz, dtype=np.uint64)
        ins[::3] = 0
        z = F.rand_elems(rng, nz)
        zr, zi = F.to_soa(z)
        orr = np.zeros(1, np.uint64)
        oi = np.zeros(1, np.uint64)
        oracle.mle_eval_base(_p(ins), _p(zr), _p(zi), nz, _p(orr), _p(oi))
        base = [(int(v) % P, 0) for v in ins]
        assert (int(orr[0]), int(oi[0])) == F.mle(base, z)


def _rand_circuit(rng, n_gates, n_wires, types=tuple(range(13))):
    gt = rng.choice(types, size=n_gates).astype(np.int32)
    i0 = rng.integers(0, n_wires, size=n_gates).astype(np.int32)
    i1 = rng.integers(0, n_wires, size=n_gates).astype(np.int32)
    z1 = F.rand_elems(rng, n_gates)
    return gt, i0, i1, z1


def _claim_bruteforce(gt, i0, i1, z1, rows, beta):
    """sum_c beta[c] sum_g z1[g] gate_g(rows[i0][c], rows[i1][c])."""
    acc = (0, 0)
    for c in range(len(beta)):
        v = (0, 0)
        for g in range(len(gt)):
            a, b = rows[i0[g]][c], rows[i1[g]][c]
            v = F.fadd(v, F.fmul(_gate_py(int(gt[g]), a, b), z1[g]))
        acc = F.fadd(acc, F.fmul(v, beta[c]))
    return acc


def _gate_py(tp, x, y):
    one = (1, 0)
    xy = F.fmul(x, y)
    table = {
        0: xy, 1: F.fadd(x, y), 2: F.fsub(x, y), 3: x, 4: y,
        5: F.fsub(F.fadd(x, y), xy), 6: F.fsub(F.fadd(x, y), F.fadd(xy, xy)),
        7: F.fsub(one, x), 8: F.fsub(one, xy), 9: F.fsub(F.fadd(one, xy), F.fadd(x, y)),
        10: F.fsub(F.fadd(one, F.fadd(xy, xy)), F.fadd(x, y)), 11: F.fmul(F.fsub(one, x), y),
        12: x,
    }
    return table[tp]


def _fold_rows(rows, r):
    return [[F.fadd(row[2 * i], F.fmul(r, F.fsub(row[2 * i + 1], row[2 * i])))
             for i in range(len(row) // 2)] for row in rows]


def test_early_claim_sumcheck_consistency(oracle):
    """Cubic through F(0), F(1), F(-1), F(2) at random r equals the claim of the
    tables folded with r: the early round's defining property."""
    rng = np.random.default_rng(16)
    n_wires, nc, n_gates = 6, 8, 9
    gt, i0, i1, z1 = _rand_circuit(rng, n_gates, n_wires)
    rows = [F.rand_elems(rng, nc) for _ in range(n_wires)]
    beta = F.rand_elems(rng, nc)
    vr = np.array([[e[0] for e in row] for row in rows], np.uint64).reshape(-1)
    vi = np.array([[e[1] for e in row] for row in rows], np.uint64).reshape(-1)
    br, bi = F.to_soa(beta)
    z1r, z1i = F.to_soa(z1)
    out8 = np.zeros(8, np.uint64)
    oracle.early_claim(n_gates, _p(gt), _p(i0), _p(i1), _p(z1r), _p(z1i), nc,
                       _p(vr), _p(vi), _p(br), _p(bi), _p(out8))
    vals = [(int(out8[2 * t]), int(out8[2 * t + 1])) for t in range(4)]
    # F(0) + F(1) = full sum over copies (pairs (2i, 2i+1) collapse)
    total = _claim_bruteforce(gt, i0, i1, z1, rows, beta)
    assert F.fadd(vals[0], vals[1]) == total
    r = F.rand_elems(rng, 1)[0]
    folded = _fold_rows(rows, r)
    bfold = _fold_rows([beta], r)[0]
    want = _claim_bruteforce(gt, i0, i1, z1, folded, bfold)
    assert F.lagrange_eval([0, 1, -1, 2], vals, r) == want


def test_sc2_bruteforce(oracle):
    rng = np.random.default_rng(17)
    for nv in (1, 3, 6):
        f = F.rand_elems(rng, 1 << nv)
        g = F.rand_elems(rng, 1 << nv)
        chal_rng = np.random.default_rng(nv)
        polys, chals, fin = oracle.sc2_prove(F.to_interleaved(f), F.to_interleaved(g),
                                              lambda k, p: F.rand_elems(chal_rng, 1)[0])
        claim = (0, 0)
        for a, b in zip(f, g):
            claim = F.fadd(claim, F.fmul(a, b))
        for k in range(nv):
            a = (int(polys[k][0]), int(polys[k][1]))
            b = (int(polys[k][2]), int(polys[k][3]))
            c = (int(polys[k][4]), int(polys[k][5]))
            s0, s1 = c, F.fadd(F.fadd(a, b), c)
            assert F.fadd(s0, s1) == claim, k
            r = (int(chals[k][0]), int(chals[k][1]))
            claim = F.fadd(F.fmul(F.fadd(F.fmul(a, r), b), r), c)
        z = [(int(x[0]), int(x[1])) for x in chals]
        fr, gr = F.mle(f, z), F.mle(g, z)
        assert (int(fin[0]), int(fin[1])) == fr and (int(fin[2]), int(fin[3])) == gr
        assert claim == F.fmul(fr, gr)


def test_root_of_unity_orders(oracle):
    g = oracle.root(62)
    assert F.fpow(g, 1 << 61) == (P - 1, 0)     # order exactly 2^62
    for k in (1, 2, 5, 20, 22):
        w = oracle.root(k)
        assert F.fpow(w, 1 << k) == (1, 0)
        assert F.fpow(w, 1 << (k - 1)) == (P - 1, 0)


def test_fft_naive_dft(oracle):
    rng = np.random.default_rng(18)
    for lc, lo in ((0, 2), (2, 2), (3, 5), (4, 6)):
        c = F.rand_elems(rng, 1 << lc)
        out = F.pairs(oracle.fft(F.to_interleaved(c), lc, lo))
        w = oracle.root(lo)
        for i in range(1 << lo):
            want = (0, 0)
            wi = F.fpow(w, i)
            for j, cj in enumerate(c):
                want = F.fadd(want, F.fmul(cj, F.fpow(wi, j)))
            assert out[i] == want, (lc, lo, i)


# ---- against the reference's own sources compiled in place ---------------------

needs_ref_native = pytest.mark.skipif(not RefNative.available(), reason="oracle/_ref not built")
needs_ref_sc = pytest.mark.skipif(not RefSumcheck.available(), reason="oracle/_ref not built")
needs_ref_rs = pytest.mark.skipif(not RefRS.available(), reason="oracle/_ref not built")


@needs_ref_native
def test_oracle_vs_reference_native(oracle):
    ref = RefNative()
    rng = np.random.default_rng(19)
    # fold + expand
    n = 64
    re, im = F.to_soa(F.rand_elems(rng, 2 * n))
    r = F.rand_elems(rng, 1)[0]
    a = [np.zeros(n, np.uint64) for _ in range(4)]
    oracle.fold(_p(re), _p(im), n, r[0], r[1], _p(a[0]), _p(a[1]))
    ref.fold(_p(re), _p(im), n, r[0], r[1], _p(a[2]), _p(a[3]))
    assert np.array_equal(a[0], a[2]) and np.array_equal(a[1], a[3])
    # beta
    z = F.rand_elems(rng, 9)
    zr, zi = F.to_soa(z)
    b = [np.zeros(512, np.uint64) for _ in range(4)]
    oracle.compute_beta(_p(zr), _p(zi), 9, 5, 7, _p(b[0]), _p(b[1]))
    ref.compute_beta(_p(zr), _p(zi), 9, 5, 7, _p(b[2]), _p(b[3]))
    assert np.array_equal(b[0], b[2]) and np.array_equal(b[1], b[3])
    # mle
    for nz in (1, 2, 3, 10, 13):
        ins = rng.integers(0, 1 << 63, size=1 << nz, dtype=np.uint64)
        zr, zi = F.to_soa(F.rand_elems(rng, nz))
        o = [np.zeros(1, np.uint64) for _ in range(4)]
        oracle.mle_eval_base(_p(ins), _p(zr), _p(zi), nz, _p(o[0]), _p(o[1]))
        ref.mle_eval_base(_p(ins), _p(zr), _p(zi), nz, _p(o[2]), _p(o[3]))
        assert (o[0][0], o[1][0]) == (o[2][0], o[3][0]), nz
    # non-session early sumcheck and the session claim / fold / finish
    n_wires, nc, n_gates = 12, 32, 40
    gt, i0, i1, z1 = _rand_circuit(rng, n_gates, n_wires)
    z1r, z1i = F.to_soa(z1)
    vr, vi = F.to_soa(F.rand_elems(rng, n_wires * nc))
    tabs = [F.to_soa(F.rand_elems(rng, n_wires * nc // 2)) for _ in range(2)]
    br, bi = F.to_soa(F.rand_elems(rng, nc))
    bf = [F.to_soa(F.rand_elems(rng, nc // 2)) for _ in range(2)]
    outs = [np.zeros(8, np.uint64), np.zeros(8, np.uint64)]
    for fn, out in ((oracle.sumcheck_early, outs[0]), (ref.sumcheck_early, outs[1])):
        fn(n_gates, _p(gt), _p(i0), _p(i1), _p(z1r), _p(z1i), nc, n_wires, _p(vr), _p(vi),
           _p(tabs[0][0]), _p(tabs[0][1]), _p(tabs[1][0]), _p(tabs[1][1]), _p(br), _p(bi),
           _p(bf[0][0]), _p(bf[0][1]), _p(bf[1][0]), _p(bf[1][1]), _p(out))
    assert np.array_equal(outs[0], outs[1])
    h = ref.early_create(n_gates, _p(gt), _p(i0), _p(i1), _p(z1r), _p(z1i), nc, n_wires,
                         _p(vr), _p(vi), _p(br), _p(bi))
    assert h
    cur = (vr.copy(), vi.copy(), br.copy(), bi.copy())
    m = nc
    while m > 1:
        o_ref, o_or = np.zeros(8, np.uint64), np.zeros(8, np.uint64)
        assert ref.early_claim(h, _p(o_ref)) == 0
        oracle.early_claim(n_gates, _p(gt), _p(i0), _p(i1), _p(z1r), _p(z1i), m,
                           _p(cur[0]), _p(cur[1]), _p(cur[2]), _p(cur[3]), _p(o_or))
        assert np.array_equal(o_ref, o_or), m
        r = F.rand_elems(rng, 1)[0]
        assert ref.early_fold(h, r[0], r[1]) == 0
        nxt = [np.zeros(n_wires * m // 2, np.uint64) for _ in range(2)] + \
              [np.zeros(m // 2, np.uint64) for _ in range(2)]
        oracle.early_fold(n_wires, m, _p(cur[0]), _p(cur[1]), _p(cur[2]), _p(cur[3]), r[0], r[1],
                          _p(nxt[0]), _p(nxt[1]), _p(nxt[2]), _p(nxt[3]))
        cur = tuple(nxt)
        m //= 2
    wr, wi = np.zeros(n_wires, np.uint64), np.zeros(n_wires, np.uint64)
    b1, b2 = np.zeros(1, np.uint64), np.zeros(1, np.uint64)
    assert ref.early_finish(h, _p(wr), _p(wi), _p(b1), _p(b2)) == 0
    ref.early_destroy(h)
    assert np.array_equal(wr, cur[0]) and np.array_equal(wi, cur[1])
    assert (b1[0], b2[0]) == (cur[2][0], cur[3][0])


@needs_ref_sc
def test_oracle_vs_reference_sumcheck(oracle):
    ref = RefSumcheck()
    rng = np.random.default_rng(20)
    for nv in (1, 2, 7, 12):
        f, g = rand_fp2(rng, 1 << nv), rand_fp2(rng, 1 << nv)
        polys, chals, fin = oracle.sc2_prove(f, g, lambda k, p: tuple(int(x) for x in rand_fp2(rng, 1)))
        rp, rfin = ref.prove(f, g, chals)
        assert np.array_equal(polys, rp) and np.array_equal(fin, rfin), nv


@needs_ref_rs
def test_oracle_vs_reference_fft(oracle):
    ref = RefRS()
    rng = np.random.default_rng(21)
    for lg in (1, 8, 20, 22, 61):
        assert oracle.root(lg) == ref.root(lg)
    for lc, lo in ((0, 0), (1, 3), (5, 7), (8, 10), (10, 12), (12, 12)):
        c = rand_fp2(rng, 1 << lc)
        assert np.array_equal(oracle.fft(c, lc, lo), ref.fft(c, lc, lo)), (lc, lo)
