"""The frame's macroblock layout on the device: YUV 4:2:0 planes to the
prover's per-copy input rows, and back.

The reference prepares a frame on the host in numpy and hands the prover and
the signer a text file, one line per macroblock
(VideoEdition/Frame_Block_Extraction.py:67-94 and :220-240; the crop edit
appends each macroblock's bit, VideoEdition/Crop.py:25-37 and :79-85), and
rebuilds the edited frame the same way (VideoEdition/Encode.py:19-42).  Here
``libtruepix_hip.so`` builds the rows on the device (``tp_yuv_pack``) and
writes planes back from edited rows (``tp_yuv_unpack``):

  - ``macroblocks(y, cb, cr, extra=None)``: host planes -> [copies][384]
    (385 with ``extra``, one byte per copy) uint8 rows;
  - ``frame(rows, width, height)``: rows -> (Y, Cb, Cr) host planes;
  - ``macroblocks_device`` / ``frame_device``: the same on torch tensors
    already on the GPU, with no host copy;
  - ``crop_bits``: the crop edit's per-macroblock bit;
  - ``edited_rows``: a frame proof's outputs as the rows ``frame`` takes.
"""
from __future__ import annotations

import ctypes

import numpy as np

from truepix_amd import native_fp2_bridge as nb

MB = 16
COPY = 384


def _planes(y, cb, cr):
    y = np.ascontiguousarray(y, dtype=np.uint8)
    cb = np.ascontiguousarray(cb, dtype=np.uint8)
    cr = np.ascontiguousarray(cr, dtype=np.uint8)
    if y.ndim != 2 or cb.shape != (y.shape[0] // 2, y.shape[1] // 2) or cr.shape != cb.shape:
        raise ValueError("planes must be Y [h][w] and Cb, Cr [h/2][w/2] (4:2:0)")
    return y, cb, cr


def copies(width: int, height: int) -> int:
    """Macroblocks (prover copies) of a width x height frame."""
    return (width // MB) * (height // MB)


def macroblocks(y, cb, cr, extra=None) -> np.ndarray:
    """Per-copy input rows of a frame: [copies][384] uint8 (Y 16x16, Cb 8x8,
    Cr 8x8 per macroblock, raster order), or [copies][385] with ``extra``
    (one byte per copy, e.g. ``crop_bits``) appended."""
    nb.require()
    y, cb, cr = _planes(y, cb, cr)
    h, w = y.shape
    n = copies(w, h)
    width = COPY + (extra is not None)
    ex = None
    if extra is not None:
        ex = np.ascontiguousarray(extra, dtype=np.uint8).reshape(-1)
        if ex.size != n:
            raise ValueError("extra needs one byte per copy (%d)" % n)
    out = np.empty((n, width), dtype=np.uint8)
    got = nb.LIB.tp_yuv_pack_host(y.ctypes.data, cb.ctypes.data, cr.ctypes.data, w, h,
                                  ex.ctypes.data if ex is not None else None, width,
                                  out.ctypes.data)
    if got < 0:
        raise RuntimeError("tp_yuv_pack_host failed: %s" % nb.last_error())
    return out


def frame(rows, width: int, height: int):
    """(Y, Cb, Cr) planes of a frame from its rows [copies][>= 384] (the
    first 384 bytes of each used); zero outside the macroblock area."""
    nb.require()
    m = np.ascontiguousarray(rows, dtype=np.uint8)
    if m.ndim != 2 or m.shape[0] != copies(width, height) or m.shape[1] < COPY:
        raise ValueError("rows must be [%d][>= 384]" % copies(width, height))
    y = np.empty((height, width), dtype=np.uint8)
    cb = np.empty((height // 2, width // 2), dtype=np.uint8)
    cr = np.empty_like(cb)
    rc = nb.LIB.tp_yuv_unpack_host(m.ctypes.data, m.shape[1], width, height, y.ctypes.data,
                                   cb.ctypes.data, cr.ctypes.data)
    if rc != 0:
        raise RuntimeError("tp_yuv_unpack_host failed: %s" % nb.last_error())
    return y, cb, cr


def macroblocks_device(y, cb, cr, extra=None, out=None, stream=None):
    """``macroblocks`` on uint8 torch tensors on the GPU (planes may be
    row-strided views); returns (or fills) a [copies][384 or 385] tensor."""
    import torch
    nb.require()
    h, w = y.shape
    n = copies(w, h)
    width = COPY + (extra is not None)
    for t in (y, cb, cr):
        if t.dtype != torch.uint8 or not t.is_cuda or t.stride(1) != 1:
            raise ValueError("planes must be uint8 CUDA tensors with unit column stride")
    if cb.stride(0) != cr.stride(0):
        raise ValueError("Cb and Cr must share a row stride")
    if out is None:
        out = torch.empty((n, width), dtype=torch.uint8, device=y.device)
    ex = extra.contiguous() if extra is not None else None
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    got = nb.LIB.tp_yuv_pack(y.data_ptr(), cb.data_ptr(), cr.data_ptr(), w, h, y.stride(0),
                             cb.stride(0), ex.data_ptr() if ex is not None else None, width,
                             out.data_ptr(), ctypes.c_void_p(st))
    if got < 0:
        raise RuntimeError("tp_yuv_pack failed: %s" % nb.last_error())
    return out


def frame_device(rows, width: int, height: int, stream=None):
    """``frame`` on a [copies][>= 384] uint8 torch tensor on the GPU; returns
    (Y, Cb, Cr) tensors."""
    import torch
    nb.require()
    if rows.dtype != torch.uint8 or not rows.is_cuda or rows.stride(1) != 1:
        raise ValueError("rows must be a uint8 CUDA tensor with unit column stride")
    y = torch.empty((height, width), dtype=torch.uint8, device=rows.device)
    cb = torch.empty((height // 2, width // 2), dtype=torch.uint8, device=rows.device)
    cr = torch.empty_like(cb)
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    rc = nb.LIB.tp_yuv_unpack(rows.data_ptr(), rows.stride(0), width, height, width, width // 2,
                              y.data_ptr(), cb.data_ptr(), cr.data_ptr(), ctypes.c_void_p(st))
    if rc != 0:
        raise RuntimeError("tp_yuv_unpack failed: %s" % nb.last_error())
    return y, cb, cr


def crop_bits(width: int, height: int, x_start: float, x_end: float, y_start: float,
              y_end: float) -> np.ndarray:
    """The crop edit's per-macroblock bit (VideoEdition/Crop.py:25-37): 1 inside
    the rectangle [x_start, x_end) x [y_start, y_end) of relative coordinates,
    the indices truncated to whole macroblocks; raster order."""
    hb, wb = int(height / 16), int(width / 16)
    b = np.zeros((hb, wb), dtype=np.uint8)
    b[int(y_start * hb):int(y_end * hb), int(x_start * wb):int(x_end * wb)] = 1
    return b.reshape(-1)


def edited_rows(outputs, n_copies: int) -> np.ndarray:
    """A frame proof's outputs (integers below 256, [padded copies][>= 384]
    or flat as ``FramePool`` keeps them with ``keep_outputs``, the copies
    padded to a power of two) as uint8 rows for
    ``frame``: the first n_copies rows, whose first 384 outputs are the
    macroblock's edited bytes in the input order (every edit circuit)."""
    o = np.asarray(outputs)
    if o.ndim == 1:   # flat, as kept: rows of the copies padded to a power of two
        pad = 1 << max(0, (n_copies - 1).bit_length())
        if o.size % pad:
            raise ValueError("outputs do not split into %d rows" % pad)
        o = o.reshape(pad, o.size // pad)
    if o.ndim != 2 or o.shape[1] < COPY or o.shape[0] < n_copies:
        raise ValueError("outputs must be [>= %d][>= 384]" % n_copies)
    o = o[:n_copies, :COPY]
    if o.size and int(o.max()) > 255:
        raise ValueError("outputs above 255 are not pixel values")
    return np.ascontiguousarray(o, dtype=np.uint8)
