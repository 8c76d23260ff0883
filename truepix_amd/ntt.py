"""Host wrapper of the Reed-Solomon NTT (tp_rs_encode).

Same mathematical result as OZ/src/infrastructure/RS_polynomial.cpp:46-130
(fast_fourier_transform): out[i] = sum_{j < 2^log_coef} c_j w^(i j) for
i < 2^log_order, w the order-2^log_order root of Orion's generator.
"""
from __future__ import annotations

import ctypes

import numpy as np

from truepix_amd import native_fp2_bridge as nb


def rs_encode_device(d_coef: int, n_cols: int, log_coef: int, log_order: int, d_out: int,
                     cols_per_chunk: int = 0, stream: int = 0) -> None:
    nb.require()
    rc = nb.LIB.tp_rs_encode(ctypes.c_void_p(d_coef), n_cols, log_coef, log_order,
                             ctypes.c_void_p(d_out), cols_per_chunk,
                             ctypes.c_void_p(stream) if stream else None)
    if rc != 0:
        raise RuntimeError("tp_rs_encode failed: %s" % nb.last_error())


def rs_encode_host(coef: np.ndarray, n_cols: int, log_coef: int, log_order: int,
                   cols_per_chunk: int = 0) -> np.ndarray:
    """coef: n_cols x 2^log_coef interleaved Fp2 (column-major); returns
    n_cols x 2^log_order interleaved evaluations."""
    import torch
    coef = np.ascontiguousarray(coef, dtype=np.uint64)
    if coef.size != n_cols * (2 << log_coef):
        raise ValueError("coefficient array has the wrong size")
    dc = torch.from_numpy(coef.view(np.int64)).to("cuda")
    do = torch.empty(n_cols * (2 << log_order), dtype=torch.int64, device="cuda")
    rs_encode_device(dc.data_ptr(), n_cols, log_coef, log_order, do.data_ptr(), cols_per_chunk,
                     torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return do.cpu().numpy().view(np.uint64)
