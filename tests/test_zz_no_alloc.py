"""GPU check of the entry point that allocates nothing,
tp_mle_eval_base_dev_noalloc: through the C ABI, on two calls in a row (the
second over a smaller input, so the scratch holds the first call's larger
partials), its value equals tp_mle_eval_base_dev's bit for bit, and a size
beyond its limit fails with an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
P = (1 << 61) - 1


def _mle(lib, fn, d_in, point):
    z = np.ascontiguousarray(np.array(point, dtype=np.uint64).reshape(-1))
    out = np.zeros(2, np.uint64)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    rc = getattr(lib, fn)(ctypes.c_void_p(d_in.data_ptr()), len(point), z.ctypes.data_as(u64p),
                          out.ctypes.data_as(u64p), None)
    return rc, (int(out[0]), int(out[1]))


def test_noalloc_mle_equals_the_allocating_entry_point(bridge, torch_cuda):
    torch = torch_cuda
    lib = bridge.LIB
    rng = np.random.default_rng(4242)
    got = []
    for n_z in (20, 12, 3):          # large, then smaller: stale partials must not leak in
        vals = rng.integers(0, P, size=1 << n_z, dtype=np.int64)
        vals[::5] = 0
        d_in = torch.from_numpy(vals).to("cuda")
        point = [(int(rng.integers(0, P)), int(rng.integers(0, P))) for _ in range(n_z)]
        rc_a, want = _mle(lib, "tp_mle_eval_base_dev", d_in, point)
        rc_b, have = _mle(lib, "tp_mle_eval_base_dev_noalloc", d_in, point)
        print("noalloc n_z", n_z, rc_a, rc_b, flush=True)
        assert rc_a == 0 and rc_b == 0, bridge.last_error()
        assert have == want, n_z
        got.append(have)
    assert len(set(got)) == len(got)
    # past the limit: an error, not a result
    d_in = torch.zeros(2, dtype=torch.int64, device="cuda")
    rc, _ = _mle(lib, "tp_mle_eval_base_dev_noalloc", d_in, [(0, 0)] * 41)
    assert rc == -1 and "n_z=41" in bridge.last_error()
