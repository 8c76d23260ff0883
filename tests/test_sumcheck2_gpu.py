"""GPU parity of the degree-2 sumcheck prover (tp_sc2_prove) against the oracle:
bit-exact round polynomials, challenges (transcript) and final values, from
tiny tables up to the 2 x 2^24 benchmark size."""
import numpy as np
import pytest

from oracle.oracle import rand_fp2
from truepix_amd import sumcheck as sc
from truepix_amd.fiatshamir import FiatShamir
from truepix_amd.extfield import _raw

pytestmark = pytest.mark.gpu


def transcript_chal(seed):
    fs = FiatShamir()
    fs.put(["c2", seed])

    def fn(k, poly6):
        fs.put([_raw(int(poly6[0]), int(poly6[1])), _raw(int(poly6[2]), int(poly6[3])),
                _raw(int(poly6[4]), int(poly6[5]))])
        r = fs.rand_ext("c2-%d" % k)
        return r.real, r.imag
    return fn


@pytest.mark.parametrize("nv", [1, 2, 3, 4, 5, 9, 10, 11, 12, 16])
def test_sc2_fixed_challenges(bridge, oracle, torch_cuda, nv):
    rng = np.random.default_rng(nv)
    f, g = rand_fp2(rng, 1 << nv), rand_fp2(rng, 1 << nv)
    chals = rand_fp2(rng, nv).reshape(nv, 2)
    want = oracle.sc2_prove(f, g, lambda k, p: (int(chals[k][0]), int(chals[k][1])))
    got = sc.prove_deg2_host(f, g, seed=0, fixed=chals)
    assert np.array_equal(got.polys, want[0])
    assert np.array_equal(got.chals, want[1])
    assert np.array_equal(got.final, want[2])


@pytest.mark.parametrize("nv,seed", [(1, 0), (2, 3), (3, 99), (4, 7), (13, 123456789)])
def test_sc2_transcript(bridge, oracle, torch_cuda, nv, seed):
    rng = np.random.default_rng(seed)
    f, g = rand_fp2(rng, 1 << nv), rand_fp2(rng, 1 << nv)
    want = oracle.sc2_prove(f, g, transcript_chal(seed))
    got = sc.prove_deg2_host(f, g, seed=seed)
    assert np.array_equal(got.polys, want[0])
    assert np.array_equal(got.chals, want[1])
    assert np.array_equal(got.final, want[2])


def test_sc2_edge_values(bridge, oracle, torch_cuda):
    P = (1 << 61) - 1
    nv = 6
    f = np.full(2 << nv, P - 1, dtype=np.uint64)
    g = np.zeros(2 << nv, dtype=np.uint64)
    g[::3] = P - 1
    want = oracle.sc2_prove(f, g, transcript_chal(5))
    got = sc.prove_deg2_host(f, g, seed=5)
    assert np.array_equal(got.polys, want[0]) and np.array_equal(got.final, want[2])


def test_sc2_full_size(bridge, oracle, torch_cuda):
    """The benchmark shape (2 x 2^24), bit-exact against the multithreaded oracle."""
    nv = 24
    f, g = sc.synthetic_tables(nv, seed=2024)
    want = oracle.sc2_prove(f, g, transcript_chal(2024), nthreads=8)
    got = sc.prove_deg2_host(f, g, seed=2024)
    assert np.array_equal(got.polys, want[0])
    assert np.array_equal(got.chals, want[1])
    assert np.array_equal(got.final, want[2])


def expected_classes(nv, groups, tail_in, mid_end, blocks):
    """The launches and bytes per class of one prove of 2^nv entries.

    Block pass (2^14 to 2^28 entries): one pass over the input for rounds 0
    to 3 ("first") and one folding it into T_4 ("fold01", reading 32 B and
    writing 2 B per input entry).  Otherwise: rounds 0 and 1 from one pass,
    then passes folding by two challenges (the first reads the input), each
    publishing the quad sums of its output while that output exceeds
    groups x tail_in entries and the pair sums after that ("pair" for all but
    the first).  Then streaming launches while the input exceeds groups x
    tail_in, the sliced launch down to mid_end entries (when the input
    exceeds tail_in), and one workgroup for the rest."""
    n = 1 << nv
    want = {c: [0, 0] for c in ("first", "fold01", "pair", "fold", "mid", "tail")}
    want["first"] = [1, 32 * n]
    if blocks and nv >= 14:
        want["fold01"] = [1, 34 * n]
        length = n // 16
    else:
        passes = [n]
        while passes[-1] // 4 > groups * tail_in:
            passes.append(passes[-1] // 4)
        want["fold01"] = [1, 40 * n]
        want["pair"] = [len(passes) - 1, sum(40 * x for x in passes[1:])]
        length = passes[-1] // 4
    big = []
    while length > groups * tail_in:
        big.append(length)
        length //= 2
    want["fold"] = [len(big), sum(48 * x for x in big)]
    if length > tail_in:
        want["mid"] = [1, 32 * length + 64 * groups]
        want["tail"] = [1, 32 * mid_end]
    else:
        want["tail"] = [1, 32 * length]
    return want


@pytest.mark.parametrize("blocks", [True, False])
def test_sc2_launch_classes(bridge, torch_cuda, blocks):
    """Per launch class, the launches and bytes of proves at sizes with and
    without streaming rounds and the sliced launch, with the block pass and
    with the quad-product passes (expected_classes gives the schedule)."""
    groups, tail_in, mid_end = sc.schedule()
    for nv in (16, 22, 24, 25):
        rng = np.random.default_rng(nv)
        f, g = rand_fp2(rng, 1 << nv), rand_fp2(rng, 1 << nv)
        sc.set_timing(True)
        sc.set_block_pass(blocks)
        try:
            sc.prove_deg2_host(f, g, seed=1)
            cls = sc.class_stats()
        finally:
            sc.set_timing(False)
            sc.set_block_pass(True)
        want = expected_classes(nv, groups, tail_in, mid_end, blocks)
        for c, (launches, nbytes) in want.items():
            assert cls[c]["launches"] == launches and cls[c]["bytes"] == nbytes, (nv, c, cls[c])
        assert all(c["ms"] > 0 for c in cls.values() if c["launches"])
        if nv == 24:   # the benchmark shape: no streaming round, one sliced launch
            assert want["fold"][0] == 0 and want["mid"][0] == 1


def test_sc2_abort_then_recover(bridge, oracle, torch_cuda):
    """A device wait that expires (forced with a zero bound) makes the prove
    fail with an error, every kernel leaving on its own; the next prove on
    the same state starts from fresh sum slots, ticket and answer slot and
    matches the oracle."""
    nv = 20
    rng = np.random.default_rng(5)
    f, g = rand_fp2(rng, 1 << nv), rand_fp2(rng, 1 << nv)
    want = oracle.sc2_prove(f, g, transcript_chal(5), nthreads=8)
    sc.set_spin_ticks(0)
    try:
        with pytest.raises(RuntimeError, match="gave up"):
            sc.prove_deg2_host(f, g, seed=5)
    finally:
        sc.set_spin_ticks(-1)
    got = sc.prove_deg2_host(f, g, seed=5)
    assert np.array_equal(got.polys, want[0])
    assert np.array_equal(got.chals, want[1])
    assert np.array_equal(got.final, want[2])


@pytest.mark.parametrize("nv", [20, 24])
def test_sc2_mid_gate_fallback(bridge, oracle, torch_cuda, nv):
    """The sliced mid launch needs all its workgroups resident; when its
    arrival gate closes before they are (forced here with a zero bound), the
    rounds go to streaming launches and the proof is unchanged (at 2^24
    after the two-round passes, reading the buffer the last pass wrote)."""
    rng = np.random.default_rng(77)
    f, g = rand_fp2(rng, 1 << nv), rand_fp2(rng, 1 << nv)
    want = oracle.sc2_prove(f, g, transcript_chal(77), nthreads=8)
    sc.set_gate_ticks(0)
    try:
        got = sc.prove_deg2_host(f, g, seed=77)
        fallbacks = sc.mid_fallbacks()
    finally:
        sc.set_gate_ticks(-1)
    assert np.array_equal(got.polys, want[0])
    assert np.array_equal(got.chals, want[1])
    assert np.array_equal(got.final, want[2])
    assert fallbacks in (0, 1)
    got2 = sc.prove_deg2_host(f, g, seed=77)
    assert sc.mid_fallbacks() == 0 and np.array_equal(got2.polys, want[0])


@pytest.mark.parametrize("nv", [3, 12, 21, 23])
def test_sc2_queued_and_serial_launches_agree(bridge, oracle, torch_cuda, nv):
    """Launches queued ahead of their transcript steps (challenges through the
    answer slot and the device relay) and launches made one after another
    give the oracle's proof, over sizes with and without streaming rounds
    and the sliced mid launch."""
    rng = np.random.default_rng(100 + nv)
    f, g = rand_fp2(rng, 1 << nv), rand_fp2(rng, 1 << nv)
    want = oracle.sc2_prove(f, g, transcript_chal(9), nthreads=8)
    try:
        for ahead in (False, True, False, True):
            sc.set_ahead(ahead)
            got = sc.prove_deg2_host(f, g, seed=9)
            assert np.array_equal(got.polys, want[0]), ahead
            assert np.array_equal(got.chals, want[1]), ahead
            assert np.array_equal(got.final, want[2]), ahead
    finally:
        sc.set_ahead(False)


@pytest.mark.parametrize("nv", [12, 13, 18])
def test_sc2_edge_values_matrix_core_pass(bridge, oracle, torch_cuda, nv):
    """Tables large enough for the first pass on the int8 matrix cores, with
    the extreme values (p - 1, 0, 1, bytes of 0x00 and 0xff, every byte at
    its signed limits) mixed in: its integer tiles, byte offsets and column
    sums must give the oracle's round polynomials exactly."""
    P = (1 << 61) - 1
    rng = np.random.default_rng(nv)
    f, g = rand_fp2(rng, 1 << nv), rand_fp2(rng, 1 << nv)
    specials = np.array([P - 1, 0, 1, 0x7f7f7f7f7f7f7f7f & P, 0x0080808080808080,
                         0x1fffffffffffffff - 1, 0x00ff00ff00ff00ff], dtype=np.uint64)
    for arr in (f, g):
        idx = rng.integers(0, arr.size, size=arr.size // 3)
        arr[idx] = specials[rng.integers(0, specials.size, size=idx.size)]
    f[: 1 << 10] = P - 1
    g[: 1 << 10] = P - 1
    want = oracle.sc2_prove(f, g, transcript_chal(11))
    sc.set_block_pass(False)   # at 2^18 the block pass would take these rounds
    try:
        got = sc.prove_deg2_host(f, g, seed=11)
    finally:
        sc.set_block_pass(True)
    assert np.array_equal(got.polys, want[0]) and np.array_equal(got.final, want[2])


@pytest.mark.parametrize("nv", [23, 25])
def test_sc2_two_round_passes(bridge, oracle, torch_cuda, nv):
    """Inputs above 2^22 entries without the block pass: round 2's pass
    publishes the nine quad sums of T_2 (rounds 2 and 3) and the passes after
    it fold by two challenges, with quad sums again while their output
    exceeds the sliced rounds' input (2^25: 2^23 -> 2^21 entries) and pair
    sums last; with the block pass the same proof; extreme values mixed in,
    bit-exact against the oracle."""
    P = (1 << 61) - 1
    rng = np.random.default_rng(300 + nv)
    f, g = rand_fp2(rng, 1 << nv), rand_fp2(rng, 1 << nv)
    specials = np.array([P - 1, 0, 1, 0x00ff00ff00ff00ff], dtype=np.uint64)
    for arr in (f, g):
        idx = rng.integers(0, arr.size, size=arr.size // 5)
        arr[idx] = specials[rng.integers(0, specials.size, size=idx.size)]
    want = oracle.sc2_prove(f, g, transcript_chal(nv), nthreads=8)
    try:
        for blocks in (False, True):
            sc.set_block_pass(blocks)
            got = sc.prove_deg2_host(f, g, seed=nv)
            assert np.array_equal(got.polys, want[0]), blocks
            assert np.array_equal(got.chals, want[1]), blocks
            assert np.array_equal(got.final, want[2]), blocks
    finally:
        sc.set_block_pass(True)


@pytest.mark.parametrize("nv", [14, 17, 20])
def test_sc2_block_pass_edge_values(bridge, oracle, torch_cuda, nv):
    """The block pass's byte tiles, offsets and column sums at its extremes:
    every byte 0x00 or 0xff in whole runs (the largest signed products), p - 1
    and 0 in long runs, mixed with uniform values; 2^14 (16 chunks, one a
    workgroup) to 2^20 (4 chunks a workgroup); rounds 0 to 3 from the block
    products, T_4 from the one fold, bit-exact against the oracle."""
    P = (1 << 61) - 1
    rng = np.random.default_rng(400 + nv)
    n = 1 << nv
    f, g = rand_fp2(rng, n), rand_fp2(rng, n)
    specials = np.array([P - 1, 0, 1, 0x1fffffffffffff00, 0x00ff00ff00ff00ff, 0x0080808080808080],
                        dtype=np.uint64)
    for arr in (f, g):
        idx = rng.integers(0, arr.size, size=arr.size // 3)
        arr[idx] = specials[rng.integers(0, specials.size, size=idx.size)]
    f[: n // 4] = P - 1   # a quarter of the blocks at the largest values
    g[: n // 4] = P - 1
    f[n // 2: n // 2 + 4096] = 0
    want = oracle.sc2_prove(f, g, transcript_chal(nv + 1), nthreads=8)
    got = sc.prove_deg2_host(f, g, seed=nv + 1)
    assert np.array_equal(got.polys, want[0])
    assert np.array_equal(got.chals, want[1])
    assert np.array_equal(got.final, want[2])
