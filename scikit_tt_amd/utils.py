"""Device counterpart of scikit_tt.utils.truncated_svd (scikit_tt/utils.py:111-158)."""
import ctypes

import numpy as np
import torch

from . import _lib

__all__ = ['truncated_svd']


def truncated_svd(matrix, threshold=0, max_rank=np.inf, rel_truncation=True):
    """SVD of a matrix with the reference's truncation rule: when threshold !=
    0 keep s/s[0] > threshold (rel_truncation) or s > threshold, then at most
    max_rank triplets.  ``matrix``: a 2-D numpy array or float64 device
    tensor; returns device tensors (u, s, v) with u (m x k), s (k), v (k x n).

    Computed on the device by libttals (Householder QR of the tall
    orientation, one-sided Jacobi SVD of R); singular values agree with
    LAPACK's to rounding, singular vectors up to sign.
    """
    from .tensor_train import device, _stream
    if isinstance(matrix, torch.Tensor):
        y = matrix.to(torch.float64)
        if not y.is_cuda:
            y = y.to(device())
    else:
        y = torch.from_numpy(np.ascontiguousarray(matrix, dtype=np.float64)).to(device())
    if y.dim() != 2:
        raise ValueError('matrix must be 2-dimensional')
    y = y.contiguous()
    m, n = y.shape
    kmax = min(m, n)
    cap = 0 if max_rank == np.inf else int(max_rank)
    if cap < 0 or (max_rank != np.inf and cap == 0):
        raise ValueError('max_rank must be a positive integer')
    dev = y.device
    U = torch.empty((m, kmax), dtype=torch.float64, device=dev)
    s = torch.empty(kmax, dtype=torch.float64, device=dev)
    V = torch.empty((kmax, n), dtype=torch.float64, device=dev)
    k = ctypes.c_int()
    ctx = _lib.context(dev.index)
    _lib.check(_lib.load().ttals_truncated_svd(ctx, _stream(dev), m, n, ctypes.c_void_p(y.data_ptr()),
                                               float(threshold), 1 if rel_truncation else 0, cap,
                                               ctypes.c_void_p(U.data_ptr()), ctypes.c_void_p(s.data_ptr()),
                                               ctypes.c_void_p(V.data_ptr()), ctypes.byref(k)), ctx, 'truncated_svd')
    k = k.value
    # the kernel packs U with leading dimension k, V as k rows
    return U.reshape(-1)[:m * k].reshape(m, k), s[:k], V.reshape(-1)[:k * n].reshape(k, n)
