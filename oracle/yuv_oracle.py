"""CPU restatement of the frame's macroblock layout.  TEST INFRASTRUCTURE ONLY.

Only tests/ and bench.py's cpu_baseline leg import this module, as the checker
or as the timed CPU baseline; the product package never imports it.

Restated from the reference, in numpy:
  - pack: VideoEdition/Frame_Block_Extraction.py:67-94 (process_macroblocks:
    macroblocks in raster order, width // 16 by height // 16, Y 16x16 and
    Cb, Cr 8x8 slices) and :220-240 (process_video_and_save_to_txt: one row
    per macroblock, Y, Cb, Cr flattened and concatenated);
  - the crop bit: VideoEdition/Crop.py:25-37 (b[y0:y1, x0:x1] = 1 with the
    indices int(fraction * blocks)) and :79-85 (appended to each row);
  - unpack: VideoEdition/Encode.py:19-42 (reconstruct_frame into zeroed
    planes).
Pinned by tests/test_yuv_cpu.py against closed forms (every packed byte is
the plane pixel the formula names), round trips and hand-checked small frames.
"""
from __future__ import annotations

import numpy as np

MB = 16


def pack(y: np.ndarray, cb: np.ndarray, cr: np.ndarray, extra=None) -> np.ndarray:
    """[copies][384] (or 385 with a per-copy byte appended) uint8 rows."""
    h, w = y.shape
    mbx, mby = w // MB, h // MB
    rows = []
    for my in range(mby):
        for mx in range(mbx):
            sx, sy = mx * MB, my * MB
            yb = y[sy:sy + MB, sx:sx + MB]
            cbb = cb[sy // 2:(sy + MB) // 2, sx // 2:(sx + MB) // 2]
            crb = cr[sy // 2:(sy + MB) // 2, sx // 2:(sx + MB) // 2]
            rows.append(np.concatenate([yb.reshape(-1), cbb.reshape(-1), crb.reshape(-1)]))
    out = np.array(rows, dtype=np.uint8).reshape(mbx * mby, 384)
    if extra is not None:
        out = np.concatenate([out, np.asarray(extra, dtype=np.uint8).reshape(-1, 1)], axis=1)
    return out


def unpack(mb: np.ndarray, width: int, height: int):
    """(Y, Cb, Cr) planes from rows [copies][>= 384]; zero outside the
    macroblock area."""
    mbx, mby = width // MB, height // MB
    y = np.zeros((height, width), dtype=np.uint8)
    cb = np.zeros((height // 2, width // 2), dtype=np.uint8)
    cr = np.zeros((height // 2, width // 2), dtype=np.uint8)
    idx = 0
    for my in range(mby):
        for mx in range(mbx):
            sx, sy = mx * MB, my * MB
            row = mb[idx]
            y[sy:sy + MB, sx:sx + MB] = row[:256].reshape(MB, MB)
            cb[sy // 2:(sy + MB) // 2, sx // 2:(sx + MB) // 2] = row[256:320].reshape(8, 8)
            cr[sy // 2:(sy + MB) // 2, sx // 2:(sx + MB) // 2] = row[320:384].reshape(8, 8)
            idx += 1
    return y, cb, cr


def crop_bits(width: int, height: int, x_start: float, x_end: float, y_start: float,
              y_end: float) -> np.ndarray:
    """The crop edit's per-macroblock bit, raster order."""
    hb, wb = int(height / 16), int(width / 16)
    b = np.zeros((hb, wb), dtype=np.uint8)
    b[int(y_start * hb):int(y_end * hb), int(x_start * wb):int(x_end * wb)] = 1
    return b.reshape(-1)
