"""GPU parity of the Reed-Solomon NTT (tp_rs_encode) against the oracle:
bit-exact full transforms at small and medium sizes, and at the benchmark
size (64 x 2^20 -> 2^22) exact evaluations at sampled points (Horner) plus a
full oracle comparison of two columns."""
import numpy as np
import pytest

from oracle.oracle import rand_fp2
from truepix_amd import ntt

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("log_coef,log_order,cols", [
    (0, 0, 1), (0, 3, 2), (1, 1, 1), (2, 4, 3), (3, 3, 1), (5, 7, 2), (7, 7, 1),
    (8, 10, 2), (10, 12, 2), (11, 13, 1), (12, 12, 1), (14, 16, 3), (17, 17, 1),
    # rates 1/16 and 1/64: 16 and 64 cosets twisted in pass A's first stage
    (10, 14, 2), (9, 15, 1),
    # 16 pass-A tiles: the XCD-ordered grid of pass A at 16 and 64 cosets
    (12, 16, 2), (13, 19, 1)])
def test_rs_encode_vs_oracle(bridge, oracle, torch_cuda, log_coef, log_order, cols):
    rng = np.random.default_rng(log_coef * 100 + log_order)
    coef = rand_fp2(rng, cols << log_coef)
    got = ntt.rs_encode_host(coef, cols, log_coef, log_order, cols_per_chunk=2)
    m = 2 << log_coef
    n = 2 << log_order
    for c in range(cols):
        want = oracle.fft(coef[c * m:(c + 1) * m], log_coef, log_order)
        assert np.array_equal(got[c * n:(c + 1) * n], want), c


def test_rs_encode_on_two_streams_and_two_orders(bridge, oracle, torch_cuda):
    """Calls alternating between two streams and two output sizes (the plan's
    tables are rebuilt at each change) give the oracle's transforms."""
    torch = torch_cuda
    rng = np.random.default_rng(77)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    cases = [(11, 13, s1), (11, 15, s2), (12, 16, s1), (11, 13, s2)]
    outs = []
    for lc, lo, st in cases:
        coef = rand_fp2(rng, 1 << lc)
        with torch.cuda.stream(st):
            d_in = torch.from_numpy(coef.view(np.int64)).to("cuda")
            d_out = torch.empty(2 << lo, dtype=torch.int64, device="cuda")
            ntt.rs_encode_device(d_in.data_ptr(), 1, lc, lo, d_out.data_ptr(), 0, st.cuda_stream)
        outs.append((coef, lc, lo, d_out, st))
    for coef, lc, lo, d_out, st in outs:
        st.synchronize()
        assert np.array_equal(d_out.cpu().numpy().view(np.uint64), oracle.fft(coef, lc, lo)), (lc, lo)


def test_rs_encode_edge_coefficients(bridge, oracle, torch_cuda):
    P = (1 << 61) - 1
    coef = np.full(2 << 10, P - 1, dtype=np.uint64)
    coef[::7] = 0
    got = ntt.rs_encode_host(coef, 1, 10, 12)
    assert np.array_equal(got, oracle.fft(coef, 10, 12))


def test_rs_encode_benchmark_shape(bridge, oracle, torch_cuda):
    """64 columns x 2^20 -> 2^22: sampled outputs of every column equal the
    polynomial evaluated at w^i; two whole columns equal the oracle."""
    cols, lc, lo = 64, 20, 22
    rng = np.random.default_rng(2022)
    coef = rand_fp2(rng, cols << lc)
    got = ntt.rs_encode_host(coef, cols, lc, lo)
    w = oracle.root(lo)
    m, n = 2 << lc, 2 << lo
    for c in range(cols):
        col = coef[c * m:(c + 1) * m]
        for i in [0, 1, (1 << lo) - 1, int(rng.integers(0, 1 << lo))]:
            want = oracle.poly_eval(col, oracle.pow(w, i))
            assert (int(got[c * n + 2 * i]), int(got[c * n + 2 * i + 1])) == want, (c, i)
    for c in (0, cols - 1):
        want = oracle.fft(coef[c * m:(c + 1) * m], lc, lo)
        assert np.array_equal(got[c * n:(c + 1) * n], want), c
