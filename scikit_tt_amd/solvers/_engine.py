"""Shared host plumbing for the fused ALS engine calls (libttals.so)."""
import collections.abc
import ctypes
import warnings

import numpy as np
import torch

from .. import _lib
from ..tensor_train import TT


class SolveInfo:
    """Per-system engine report: status, sweeps, Krylov iterations, eigenvalue."""

    __slots__ = ('status', 'sweeps', 'unconverged', 'max_iters', 'krylov_iters', 'local_solves', 'eigenvalue',
                 'worst_residual', 'flops', 'device_ms', 'bytes', '_r')
    _CYCLE_NAMES = ('total', 'matvec', 'gram_schmidt', 'tridiag', 'qr', 'stacks', 'setup')

    def __init__(self, r):
        self.status = int(r.status)
        self.sweeps = int(r.sweeps)
        self.unconverged = int(r.unconverged)
        self.krylov_iters = int(r.krylov_iters)
        self.local_solves = int(r.local_solves)
        self.max_iters = int(r.max_iters)
        self.eigenvalue = float(r.lambda_)
        self.worst_residual = float(r.worst_residual)
        self.flops = float(r.flops)
        self.device_ms = float(r.device_ms)
        self.bytes = float(r.bytes)
        self._r = r  # keeps the result record for the (rarely read) cycle counters

    @property
    def cycles(self):
        """Per-phase cycle counters (non-zero in a -DTTALS_PROFILE=1 build)."""
        return {k: float(v) for k, v in zip(self._CYCLE_NAMES, self._r.cycles)}

    def __repr__(self):
        names = [k for k in self.__slots__ if k != '_r'] + ['cycles']
        return 'SolveInfo(%s)' % ', '.join('%s=%r' % (k, getattr(self, k)) for k in names)


RESULT_DTYPE = np.dtype([('status', '<i4'), ('sweeps', '<i4'), ('unconverged', '<i4'), ('max_iters', '<i4'),
                         ('krylov_iters', '<i8'), ('local_solves', '<i8'), ('eigenvalue', '<f8'),
                         ('worst_residual', '<f8'), ('flops', '<f8'), ('cycles', '<f8', (7,)),
                         ('device_ms', '<f8'), ('bytes', '<f8')])
assert RESULT_DTYPE.itemsize == ctypes.sizeof(_lib.Result)


class Infos(collections.abc.Sequence):
    """Per-system reports of one batched call: ``infos[s]`` is a SolveInfo
    (made on access); ``infos.array`` views every report as one numpy record
    array (fields status, sweeps, ..., eigenvalue, flops), without
    per-system Python objects."""

    def __init__(self, res):
        self._res = res
        self.array = np.frombuffer(res, dtype=RESULT_DTYPE)

    def __len__(self):
        return len(self._res)

    def __getitem__(self, s):
        if isinstance(s, slice):
            return [SolveInfo(self._res[k]) for k in range(*s.indices(len(self)))]
        if s < 0:
            s += len(self)
        return SolveInfo(self._res[s])


class Solutions(collections.abc.Sequence):
    """Solution trains of one batched call, each made on access (views of
    one device buffer): a call allocates O(1) Python objects however many
    systems it solves."""

    def __init__(self, cores, batch, initial_guesses):
        self._cores, self._batch, self._x0 = cores, batch, initial_guesses

    def __len__(self):
        return len(self._x0)

    def __getitem__(self, s):
        if isinstance(s, slice):
            return [self[k] for k in range(*s.indices(len(self)))]
        if s < 0:
            s += len(self)
        return finish(self._cores, s, self._batch, self._x0[s])


class EigenResults(collections.abc.Sequence):
    """(eigenvalue, TT, iterations) per system of one batched evp call, made
    on access; ``eigenvalues`` is a numpy array of all of them."""

    def __init__(self, sols, infos):
        self._sols, self._infos = sols, infos
        self.eigenvalues = infos.array['eigenvalue'].copy()

    def __len__(self):
        return len(self._sols)

    def __getitem__(self, s):
        if isinstance(s, slice):
            return [self[k] for k in range(*s.indices(len(self)))]
        if s < 0:
            s += len(self)
        return (float(self.eigenvalues[s]), self._sols[s], int(self._infos.array['sweeps'][s]))


def raise_first(infos, what, count):
    """raise_status for the first system whose status is not 0 (vectorised scan)."""
    bad = np.nonzero(infos.array['status'] != 0)[0]
    if len(bad):
        s = int(bad[0])
        raise_status(infos[s], '%s system %d' % (what, s) if count > 1 else what)


def check_tts(*groups):
    for g in groups:
        for t in g:
            if not isinstance(t, TT):
                raise TypeError('Unsupported argument.')


def descs(tts):
    """ttals_tt[] for the trains: (DescBatch, ctypes array to pass); the array owns all descriptor memory."""
    b = _lib.DescBatch.from_tts(tts)
    return b, b.ptr


def outputs(initial_guesses):
    """Output cores with the capacity of the initial guesses (ranks never grow).

    One device allocation holds every output core of the batch; system s
    gets the offsets ``slots[s]`` (in doubles) into it.  Returns
    ((buffer, slots), descriptor owner, descriptor pointer).
    """
    dev = initial_guesses[0].device
    x0 = initial_guesses[0]
    n = len(initial_guesses)

    def layout(x):
        r = np.asarray(x.ranks, dtype=np.int64)
        sizes = r[:-1] * np.asarray(x.row_dims, dtype=np.int64) * np.asarray(x.col_dims, dtype=np.int64) * r[1:]
        return np.concatenate(([0], np.cumsum(sizes)[:-1])), int(sizes.sum())

    if all(x.row_dims == x0.row_dims and x.col_dims == x0.col_dims and x.ranks == x0.ranks for x in initial_guesses):
        offs, per = layout(x0)  # one structure (the usual batch): slots are strided copies
        slots = np.arange(n, dtype=np.int64)[:, None] * per + offs[None, :]
        total = n * per
        orders, rows, cols, ranks = [x0.order] * n, np.tile(x0.row_dims, n), np.tile(x0.col_dims, n), \
            np.tile(x0.ranks, n)
        flat = slots.ravel()
    else:
        orders, rows, cols, ranks, slots = [], [], [], [], []
        total = 0
        for x in initial_guesses:
            offs, per = layout(x)
            slots.append(total + offs)
            total += per
            orders.append(x.order)
            rows += x.row_dims
            cols += x.col_dims
            ranks += x.ranks
        flat = np.concatenate(slots)
    buf = torch.empty(max(total, 1), dtype=torch.float64, device=dev)
    ptrs = np.uint64(buf.data_ptr()) + np.uint64(8) * flat.astype(np.uint64)
    b = _lib.DescBatch(orders, rows, cols, ranks, ptrs)
    return (buf, slots), b, b.ptr


def finish(cores, s, batch, x0):
    """System s's solution: the engine's in-place cores with their final ranks
    (each core at the start of its slot), as views of the batch buffer."""
    buf, slots = cores
    return TT._from_flat(buf, slots[s], x0.row_dims, x0.col_dims, batch.ranks_of(s))


def raise_status(info, what):
    """Map an engine status to the reference's exception (or a warning)."""
    st = info.status
    if st == _lib.S_SINGULAR:
        raise np.linalg.LinAlgError('Singular matrix (%s)' % what)
    if st == _lib.S_NONFINITE:
        raise np.linalg.LinAlgError('non-finite local solution (%s)' % what)
    if st == _lib.S_UNSUPPORTED:
        raise NotImplementedError('%s: the eigenvalue nearest sigma lies inside the spectrum of a local problem '
                                  'larger than the Krylov space' % what)
    if st == _lib.S_NONSYM:
        raise NotImplementedError('%s: the operator is not symmetric; local eigensolves of non-symmetric '
                                  'operators (the reference\'s dgeev path) are not implemented' % what)
    if st == _lib.S_DEADLINE:
        raise TimeoutError('%s: the time budget ran out after %d local solves' % (what, info.local_solves))
    if st == _lib.S_NOCONV:
        warnings.warn('%s: %d local Krylov solve(s) stopped at maxit' % (what, info.unconverged), RuntimeWarning)
