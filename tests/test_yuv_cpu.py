"""CPU checks of the macroblock layout oracle (oracle/yuv_oracle.py,
restating VideoEdition/Frame_Block_Extraction.py:67-94 and :220-240,
Crop.py:25-37, Encode.py:19-42) against values worked out by hand, round
trips and the frame sizes the benchmarks use."""
import numpy as np

from oracle import yuv_oracle as yo


def coord_planes(w, h):
    """Planes whose pixels are simple functions of their coordinates."""
    r, c = np.mgrid[0:h, 0:w]
    y = ((r + 2 * c) % 256).astype(np.uint8)
    r2, c2 = np.mgrid[0:h // 2, 0:w // 2]
    cb = ((3 * r2 + c2 + 100) % 256).astype(np.uint8)
    cr = ((r2 + 5 * c2 + 7) % 256).astype(np.uint8)
    return y, cb, cr


def test_pack_hand_values():
    """A 48 x 32 frame: 3 x 2 macroblocks; entries worked out by hand from
    Y = r + 2c, Cb = 3r + c + 100, Cr = r + 5c + 7 (a transposed block, a
    swapped plane or a wrong macroblock order changes them)."""
    y, cb, cr = coord_planes(48, 32)
    m = yo.pack(y, cb, cr)
    assert m.shape == (6, 384)
    assert m[0, 0] == 0                      # Y[0][0]
    assert m[1, 0] == 32                     # copy 1 = (mx 1, my 0): Y[0][16]
    assert m[1, 17] == 1 + 2 * 17            # Y[1][17]
    assert m[3, 255] == 31 + 2 * 15          # copy 3 = (mx 0, my 1): Y[31][15]
    assert m[5, 0] == 16 + 2 * 32            # copy 5 = (mx 2, my 1): Y[16][32]
    assert m[0, 256] == 100                  # Cb[0][0]
    assert m[4, 256 + 9] == 3 * 9 + 9 + 100  # copy 4 = (mx 1, my 1): Cb[8 + 1][8 + 1]
    assert m[2, 320] == 5 * 16 + 7           # copy 2 = (mx 2, my 0): Cr[0][16]
    assert m[5, 383] == 15 + 5 * 23 + 7      # Cr[8 + 7][16 + 7]


def test_partial_macroblocks_dropped_and_sizes():
    for (w, h, n) in ((1920, 1080, 8040), (1280, 720, 3600), (720, 480, 1350), (320, 240, 300),
                      (64, 64, 16), (50, 34, 6)):
        y, cb, cr = coord_planes(w, h)
        assert yo.pack(y, cb, cr).shape == (n, 384), (w, h)


def test_round_trips():
    rng = np.random.default_rng(3)
    for (w, h) in ((64, 48), (50, 34), (320, 240)):
        y = rng.integers(0, 256, (h, w), dtype=np.uint8)
        cb = rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8)
        cr = rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8)
        m = yo.pack(y, cb, cr)
        y2, cb2, cr2 = yo.unpack(m, w, h)
        hy, wy = (h // 16) * 16, (w // 16) * 16
        assert np.array_equal(y2[:hy, :wy], y[:hy, :wy]) and not y2[hy:].any() and not y2[:, wy:].any()
        assert np.array_equal(cb2[:hy // 2, :wy // 2], cb[:hy // 2, :wy // 2])
        assert np.array_equal(cr2[:hy // 2, :wy // 2], cr[:hy // 2, :wy // 2])
        assert not cb2[hy // 2:].any() and not cr2[:, wy // 2:].any()
        assert np.array_equal(yo.pack(y2, cb2, cr2), m)


def test_crop_bits_hand_values():
    """720 x 480: 45 x 30 macroblocks.  x in [0.25, 0.75) -> columns 11..32,
    y in [0.5, 1.0) -> rows 15..29: 22 x 15 ones."""
    b = yo.crop_bits(720, 480, 0.25, 0.75, 0.5, 1.0)
    assert b.shape == (1350,) and int(b.sum()) == 22 * 15
    assert b[15 * 45 + 11] == 1 and b[15 * 45 + 10] == 0 and b[14 * 45 + 11] == 0
    assert b[29 * 45 + 32] == 1 and b[29 * 45 + 33] == 0
    y, cb, cr = coord_planes(720, 480)
    m = yo.pack(y, cb, cr, extra=b)
    assert m.shape == (1350, 385) and np.array_equal(m[:, 384], b)


def test_product_crop_bits_match():
    from truepix_amd import yuv
    for args in ((720, 480, 0.25, 0.75, 0.5, 1.0), (1920, 1080, 0.1, 0.9, 0.0, 0.33),
                 (64, 64, 0.0, 1.0, 0.0, 1.0)):
        assert np.array_equal(yuv.crop_bits(*args), yo.crop_bits(*args)), args
