"""GPU parity of the 14 drop-in native_fp2 entry points against the oracle,
through the C ABI (ctypes), bit-exact, on seeded inputs and edge shapes."""
import ctypes

import numpy as np
import pytest

from oracle.oracle import _p
from tests import fieldpy as F

pytestmark = pytest.mark.gpu
P = F.P


def rand_soa(rng, n):
    x = rng.integers(0, P, size=(2, n), dtype=np.int64).astype(np.uint64)
    return np.ascontiguousarray(x[0]), np.ascontiguousarray(x[1])


@pytest.mark.parametrize("n_out", [1, 3, 255, 256, 4097, 1 << 20])
def test_fold(bridge, oracle, n_out):
    rng = np.random.default_rng(n_out)
    re, im = rand_soa(rng, 2 * n_out)
    re[:4] = [0, P - 1, 1, 0][: min(4, re.size)]
    r = rand_soa(rng, 1)
    got = [np.zeros(n_out, np.uint64) for _ in range(2)]
    want = [np.zeros(n_out, np.uint64) for _ in range(2)]
    bridge.LIB.native_fp2_fold(_p(re), _p(im), n_out, int(r[0][0]), int(r[1][0]), _p(got[0]), _p(got[1]))
    bridge.check("fold")
    oracle.fold(_p(re), _p(im), n_out, int(r[0][0]), int(r[1][0]), _p(want[0]), _p(want[1]))
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


@pytest.mark.parametrize("n_src,nc", [(1, 1), (5, 3), (512, 4), (3, 1000)])
def test_expand(bridge, oracle, n_src, nc):
    rng = np.random.default_rng(n_src * 7 + nc)
    re, im = rand_soa(rng, n_src)
    got = [np.zeros(n_src * nc, np.uint64) for _ in range(2)]
    bridge.LIB.native_fp2_expand(_p(re), _p(im), n_src, nc, _p(got[0]), _p(got[1]))
    bridge.check("expand")
    assert np.array_equal(got[0], np.repeat(re, nc)) and np.array_equal(got[1], np.repeat(im, nc))


@pytest.mark.parametrize("nz", [1, 2, 7, 8, 9, 11, 16, 20])
def test_compute_beta(bridge, oracle, nz):
    rng = np.random.default_rng(nz)
    zr, zi = rand_soa(rng, nz)
    init = (int(rng.integers(0, P)), int(rng.integers(0, P)))
    n = 1 << nz
    got = [np.zeros(n, np.uint64) for _ in range(2)]
    want = [np.zeros(n, np.uint64) for _ in range(2)]
    bridge.LIB.native_fp2_compute_beta(_p(zr), _p(zi), nz, init[0], init[1], _p(got[0]), _p(got[1]))
    bridge.check("compute_beta")
    oracle.compute_beta(_p(zr), _p(zi), nz, init[0], init[1], _p(want[0]), _p(want[1]))
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


@pytest.mark.parametrize("nz", [0, 1, 2, 9, 10, 11, 20])
def test_mle_eval_base(bridge, oracle, nz):
    rng = np.random.default_rng(50 + nz)
    ins = rng.integers(0, 1 << 63, size=1 << nz, dtype=np.uint64)
    ins[::5] = 0
    ins[1::7] = np.uint64((1 << 64) - 1)    # unreduced inputs exercise reduce64
    zr, zi = rand_soa(rng, max(nz, 1))
    got = [np.zeros(1, np.uint64) for _ in range(2)]
    want = [np.zeros(1, np.uint64) for _ in range(2)]
    bridge.LIB.native_fp2_mle_eval_base(_p(ins), _p(zr), _p(zi), nz, _p(got[0]), _p(got[1]))
    bridge.check("mle_eval_base")
    oracle.mle_eval_base(_p(ins), _p(zr), _p(zi), nz, _p(want[0]), _p(want[1]))
    assert (got[0][0], got[1][0]) == (want[0][0], want[1][0])


def _circuit(rng, n_gates, n_wires):
    gt = rng.integers(0, 13, size=n_gates).astype(np.int32)
    i0 = rng.integers(0, n_wires, size=n_gates).astype(np.int32)
    i1 = rng.integers(0, n_wires, size=n_gates).astype(np.int32)
    z1r, z1i = rand_soa(rng, n_gates)
    return gt, i0, i1, z1r, z1i


@pytest.mark.parametrize("n_gates,n_wires,nc", [(1, 1, 2), (13, 5, 4), (40, 16, 128), (384, 512, 64)])
def test_sumcheck_early_tables(bridge, oracle, n_gates, n_wires, nc):
    rng = np.random.default_rng(n_gates + nc)
    gt, i0, i1, z1r, z1i = _circuit(rng, n_gates, n_wires)
    vr, vi = rand_soa(rng, n_wires * nc)
    f0 = rand_soa(rng, n_wires * nc // 2)
    f1 = rand_soa(rng, n_wires * nc // 2)
    br, bi = rand_soa(rng, nc)
    b0 = rand_soa(rng, nc // 2)
    b1 = rand_soa(rng, nc // 2)
    args = [n_gates, _p(gt), _p(i0), _p(i1), _p(z1r), _p(z1i), nc, n_wires, _p(vr), _p(vi),
            _p(f0[0]), _p(f0[1]), _p(f1[0]), _p(f1[1]), _p(br), _p(bi), _p(b0[0]), _p(b0[1]),
            _p(b1[0]), _p(b1[1])]
    got, want = np.zeros(8, np.uint64), np.zeros(8, np.uint64)
    bridge.LIB.native_fp2_sumcheck_early(*args, _p(got))
    bridge.check("sumcheck_early")
    oracle.sumcheck_early(*args, _p(want))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n_gates,n_wires,nc", [(1, 1, 2), (9, 6, 8), (384, 512, 2048), (100, 64, 1 << 13)])
def test_early_session(bridge, oracle, n_gates, n_wires, nc):
    rng = np.random.default_rng(1000 + nc)
    gt, i0, i1, z1r, z1i = _circuit(rng, n_gates, n_wires)
    vr, vi = rand_soa(rng, n_wires * nc)
    br, bi = rand_soa(rng, nc)
    lib = bridge.LIB
    h = lib.native_fp2_early_create(n_gates, _p(gt), _p(i0), _p(i1), _p(z1r), _p(z1i), nc, n_wires,
                                    _p(vr), _p(vi), _p(br), _p(bi))
    assert h, bridge.last_error()
    try:
        assert lib.native_fp2_early_n_copies(h) == nc and lib.native_fp2_early_n_wires(h) == n_wires
        cur = [vr, vi, br, bi]
        m = nc
        while m > 1:
            got, want = np.zeros(8, np.uint64), np.zeros(8, np.uint64)
            assert lib.native_fp2_early_claim(h, _p(got)) == 0, bridge.last_error()
            oracle.early_claim(n_gates, _p(gt), _p(i0), _p(i1), _p(z1r), _p(z1i), m,
                               _p(cur[0]), _p(cur[1]), _p(cur[2]), _p(cur[3]), _p(want))
            assert np.array_equal(got, want), m
            r = rand_soa(rng, 1)
            assert lib.native_fp2_early_fold(h, int(r[0][0]), int(r[1][0])) == 0
            nxt = [np.zeros(n_wires * m // 2, np.uint64) for _ in range(2)] + \
                  [np.zeros(m // 2, np.uint64) for _ in range(2)]
            oracle.early_fold(n_wires, m, *[_p(c) for c in cur], int(r[0][0]), int(r[1][0]),
                              *[_p(x) for x in nxt])
            cur = nxt
            m //= 2
        assert lib.native_fp2_early_n_copies(h) == 1
        # folded to one copy: further claims / folds are errors, not crashes
        assert lib.native_fp2_early_claim(h, _p(np.zeros(8, np.uint64))) == -1
        assert lib.native_fp2_early_fold(h, 1, 0) == -1
        wr, wi = np.zeros(n_wires, np.uint64), np.zeros(n_wires, np.uint64)
        b = [np.zeros(1, np.uint64), np.zeros(1, np.uint64)]
        assert lib.native_fp2_early_finish(h, _p(wr), _p(wi), _p(b[0]), _p(b[1])) == 0
        assert np.array_equal(wr, cur[0]) and np.array_equal(wi, cur[1])
        assert (b[0][0], b[1][0]) == (cur[2][0], cur[3][0])
    finally:
        lib.native_fp2_early_destroy(h)


def test_early_finish_before_full_fold(bridge):
    """finish() with copies left reads element 0 of every row (strided gather)."""
    rng = np.random.default_rng(7)
    n_gates, n_wires, nc = 3, 5, 16
    gt, i0, i1, z1r, z1i = _circuit(rng, n_gates, n_wires)
    vr, vi = rand_soa(rng, n_wires * nc)
    br, bi = rand_soa(rng, nc)
    lib = bridge.LIB
    h = lib.native_fp2_early_create(n_gates, _p(gt), _p(i0), _p(i1), _p(z1r), _p(z1i), nc, n_wires,
                                    _p(vr), _p(vi), _p(br), _p(bi))
    assert h
    wr, wi = np.zeros(n_wires, np.uint64), np.zeros(n_wires, np.uint64)
    b = [np.zeros(1, np.uint64), np.zeros(1, np.uint64)]
    assert lib.native_fp2_early_finish(h, _p(wr), _p(wi), _p(b[0]), _p(b[1])) == 0
    lib.native_fp2_early_destroy(h)
    assert np.array_equal(wr, vr[::nc]) and np.array_equal(wi, vi[::nc]) and b[0][0] == br[0]


@pytest.mark.parametrize("n_gates,n_wires,nc", [(0, 4, 4), (3, 0, 4), (3, 4, 1), (3, 4, 6)])
def test_early_create_rejects_bad_shapes(bridge, n_gates, n_wires, nc):
    z = np.zeros(max(1, n_wires * nc), np.uint64)
    g = np.zeros(max(1, n_gates), np.int32)
    h = bridge.LIB.native_fp2_early_create(n_gates, _p(g), _p(g), _p(g), _p(z), _p(z), nc, n_wires,
                                           _p(z), _p(z), _p(z), _p(z))
    assert not h


def test_null_handles(bridge):
    lib = bridge.LIB
    out = np.zeros(8, np.uint64)
    assert lib.native_fp2_early_claim(None, _p(out)) == -1
    assert lib.native_fp2_early_fold(None, 1, 2) == -1
    assert lib.native_fp2_early_n_copies(None) == 0
    lib.native_fp2_early_destroy(None)


def test_bridge_functions_match_reference_shapes(bridge, oracle):
    """The Python bridge API (fold_values / compute_beta / mle_eval_base / session)."""
    from truepix_amd.extfield import Fp2
    rng = np.random.default_rng(3)
    vals = [Fp2(int(a), int(b)) for a, b in zip(*rand_soa(rng, 16))]
    r = Fp2(12345, 678)
    out, scratch = bridge.fold_values(vals, r, 4)
    assert len(scratch) == 8 and len(out) == 32
    for i in range(8):
        a, b = vals[2 * i], vals[2 * i + 1]
        assert scratch[i] == a + r * (b - a)
        assert all(out[4 * i + j] == scratch[i] for j in range(4))
    z = [Fp2(3, 4), Fp2(5, 6), 7]
    beta = bridge.compute_beta(z, 2)
    assert len(beta) == 8
    for x in range(8):
        want = Fp2(2, 0)
        for j in range(3):
            want = want * (z[j] if (x >> j) & 1 else 1 - z[j])
        assert beta[x] == want
    ins = list(range(1, 9))
    v = bridge.mle_eval_base(ins, z)
    assert v == sum((beta[x] * ins[x] for x in range(8)), Fp2(0, 0)) * Fp2(2, 0).inverse()
