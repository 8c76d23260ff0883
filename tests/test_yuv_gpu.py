"""GPU parity of the macroblock layout (tp_yuv_pack / tp_yuv_unpack) against
the oracle, bit-exact, at the benchmark frame sizes and odd ones; the prover
and the signer fed from planes give the proofs and roots of the same rows."""
import numpy as np
import pytest

from oracle import yuv_oracle as yo
from truepix_amd import circuits, frames, orion, yuv
from truepix_amd.fp2_link import commit_binding, make_manifest

pytestmark = pytest.mark.gpu
P = (1 << 61) - 1
SIZES = ((16, 16), (50, 34), (64, 48), (320, 240), (720, 480), (1280, 720), (1920, 1080))


def planes(w, h, seed):
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 256, (h, w), dtype=np.uint8),
            rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8),
            rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8))


@pytest.mark.parametrize("w,h", SIZES)
def test_pack_unpack_match_oracle(bridge, w, h):
    y, cb, cr = planes(w, h, w + h)
    m = yuv.macroblocks(y, cb, cr)
    assert np.array_equal(m, yo.pack(y, cb, cr)), (w, h)
    bits = yo.crop_bits(w, h, 0.2, 0.7, 0.1, 0.9)
    assert np.array_equal(yuv.macroblocks(y, cb, cr, extra=bits), yo.pack(y, cb, cr, extra=bits))
    edited = np.random.default_rng(1).integers(0, 256, m.shape, dtype=np.uint8)
    got = yuv.frame(edited, w, h)
    want = yo.unpack(edited, w, h)
    for a, b in zip(got, want):
        assert np.array_equal(a, b), (w, h)


def test_device_tensors_with_row_strides(bridge, torch_cuda):
    import torch
    w, h = 720, 480
    y, cb, cr = planes(w, h, 9)
    # planes as views into wider buffers (row stride > width)
    by = torch.zeros((h, w + 64), dtype=torch.uint8, device="cuda")
    bc = torch.zeros((h // 2, w // 2 + 32), dtype=torch.uint8, device="cuda")
    br = torch.zeros((h // 2, w // 2 + 32), dtype=torch.uint8, device="cuda")
    by[:, :w] = torch.from_numpy(y).cuda()
    bc[:, :w // 2] = torch.from_numpy(cb).cuda()
    br[:, :w // 2] = torch.from_numpy(cr).cuda()
    m = yuv.macroblocks_device(by[:, :w], bc[:, :w // 2], br[:, :w // 2])
    torch.cuda.synchronize()
    assert np.array_equal(m.cpu().numpy(), yo.pack(y, cb, cr))
    # rows with a wider stride (a prover's [copies][512] output block)
    wide = torch.zeros((m.shape[0], 512), dtype=torch.uint8, device="cuda")
    wide[:, :384] = m
    y2, cb2, cr2 = yuv.frame_device(wide[:, :384], w, h)
    torch.cuda.synchronize()
    want = yo.unpack(yo.pack(y, cb, cr), w, h)
    assert np.array_equal(y2.cpu().numpy(), want[0])
    assert np.array_equal(cb2.cpu().numpy(), want[1]) and np.array_equal(cr2.cpu().numpy(), want[2])


def test_bad_frames_rejected(bridge):
    y, cb, cr = planes(64, 48, 1)
    with pytest.raises(ValueError):
        yuv.macroblocks(y, cb[:, :10], cr)
    with pytest.raises(RuntimeError):
        yuv.macroblocks(y[:8, :8], cb[:4, :4], cr[:4, :4])   # below one macroblock
    with pytest.raises(ValueError):
        yuv.macroblocks(y, cb, cr, extra=np.zeros(5, np.uint8))


@pytest.fixture(scope="module")
def circs(bridge):
    c = {n: frames.FrameCircuit(circuits.circuit(n)) for n in ("gray", "crop")}
    yield c
    for v in c.values():
        v.close()


def test_prover_from_planes_and_edited_frame(circs):
    """A 320 x 240 frame proved from its planes gives the proof of its rows;
    the proof's outputs written back as planes are the gray edit (Y kept,
    Cb = Cr = 128 over the macroblock area; VideoEdition/Gray.py:7-10)."""
    w, h = 320, 240
    y, cb, cr = planes(w, h, 5)
    rows = yo.pack(y, cb, cr)
    rng = np.random.default_rng(2)
    vole = tuple(rng.integers(1, P, size=500, dtype=np.int64).astype(np.uint64) for _ in range(4))
    binding = commit_binding(make_manifest(rows.shape[0], seed=4))
    with frames.FramePool(1) as pool:
        a = pool.result(pool.submit(circs["gray"], rows, 777, vole, binding, keep_outputs=True))
        b = pool.result(pool.submit_yuv(circs["gray"], y, cb, cr, 777, vole, binding,
                                        keep_outputs=True))
    assert a.proof == b.proof
    gy, gcb, gcr = yuv.frame(yuv.edited_rows(b.outputs, rows.shape[0]), w, h)
    assert np.array_equal(gy, y) and (gcb == 128).all() and (gcr == 128).all()


def test_crop_prover_from_planes(circs):
    w, h = 64, 48
    y, cb, cr = planes(w, h, 6)
    bits = yuv.crop_bits(w, h, 0.3, 0.8, 0.0, 0.6)
    rows = yo.pack(y, cb, cr, extra=bits)
    rng = np.random.default_rng(3)
    vole = tuple(rng.integers(1, P, size=500, dtype=np.int64).astype(np.uint64) for _ in range(4))
    binding = commit_binding(make_manifest(rows.shape[0], seed=5))
    with frames.FramePool(1) as pool:
        a = pool.result(pool.submit(circs["crop"], rows, 778, vole, binding, keep_outputs=True))
        b = pool.result(pool.submit_yuv(circs["crop"], y, cb, cr, 778, vole, binding, extra=bits,
                                        keep_outputs=True))
    assert a.proof == b.proof
    out = yuv.edited_rows(b.outputs, rows.shape[0])
    assert np.array_equal(out, rows[:, :384] * bits[:, None])   # VideoEdition/Crop.py:10-17


def test_signer_commit_from_planes(bridge):
    w, h = 720, 480
    y, cb, cr = planes(w, h, 8)
    assert orion.commit_yuv(y, cb, cr, seed=99) == orion.commit(yo.pack(y, cb, cr), seed=99)


def test_device_rows_in_both_drivers_and_copied_at_submit(circs):
    """Rows built on the device from planes go to the prover without leaving
    the GPU (tp_frame_job.inputs_on_device), in a stream worker and inside a
    lockstep batch mixed with host-row frames; each proof equals the proof of
    the same rows given from the host, and overwriting the device rows right
    after submit changes nothing (the library copied them)."""
    import torch
    w, h = 160, 96
    specs, want = [], []
    rng = np.random.default_rng(9)
    vole = tuple(rng.integers(1, P, size=500, dtype=np.int64).astype(np.uint64) for _ in range(4))
    with frames.FramePool(1) as pool:
        for k in range(4):
            y, cb, cr = planes(w, h, 20 + k)
            rows = yo.pack(y, cb, cr)
            binding = commit_binding(make_manifest(rows.shape[0], seed=10 + k))
            want.append(pool.result(pool.submit(circs["gray"], rows, 900 + k, vole, binding)).proof)
            if k % 2:
                specs.append((circs["gray"], rows, 900 + k, vole, binding))
            else:
                d = yuv.macroblocks_device(torch.from_numpy(y).cuda(), torch.from_numpy(cb).cuda(),
                                           torch.from_numpy(cr).cuda())
                torch.cuda.synchronize()
                specs.append((circs["gray"], d, 900 + k, vole, binding))
        d = specs[0][1]
        jid = pool.submit(*specs[0])
        d.fill_(0)                      # after submit: the job holds its own copy
        torch.cuda.synchronize()
        assert pool.result(jid).proof == want[0]
    with frames.FramePool(1, batched=True) as bp:
        got = [bp.result(j).proof for j in bp.submit_many(specs)]
    assert got[1:] == want[1:]
    y, cb, cr = planes(w, h, 20)
    binding = commit_binding(make_manifest(yo.pack(y, cb, cr).shape[0], seed=10))
    with frames.FramePool(1, batched=True) as bp:
        tplanes = [torch.from_numpy(a).cuda() for a in (y, cb, cr)]
        assert bp.result(bp.submit_yuv(circs["gray"], *tplanes, 900, vole, binding)).proof == want[0]
