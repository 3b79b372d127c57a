"""Device-resident tensor trains (drop-in for scikit_tt.tensor_train).

``TT`` mirrors the attributes of scikit_tt/tensor_train.py:10-267 (``order``,
``row_dims``, ``col_dims``, ``ranks``, ``cores``) but keeps every core as a
float64 HIP device tensor.  ``t.cores[i]`` returns a read-only host numpy
copy and ``t.cores[i] = arr`` uploads; the device tensors are
``t.device_cores``.

All arithmetic on the ALS path runs in libttals.so (hand-written HIP):
``dot``/``@`` (core products), ``+``/``-``/``*`` (block embedding), ``norm``
(QR sweep) and ``ortho_right``/``ortho_left`` (Householder QR sweeps).
PyTorch only provides device memory and the current stream.
"""
import ctypes
import numbers

import numpy as np
import torch

from . import _lib

__all__ = ['TT', 'zeros', 'ones', 'eye', 'rand', 'residual_error', 'residual_norms', 'device', 'from_host_batch']


def device():
    """The HIP device used for new tensor trains (torch's current device)."""
    if not torch.cuda.is_available():
        raise RuntimeError('scikit_tt_amd needs a HIP device (torch.cuda.is_available() is False)')
    return torch.device('cuda', torch.cuda.current_device())


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _as_device_core(c, dev):
    if isinstance(c, torch.Tensor):
        if c.dtype != torch.float64:
            if c.is_complex():
                raise NotImplementedError('complex tensor trains are not supported')
            c = c.to(torch.float64)
        return c.to(dev).contiguous()
    arr = np.asarray(c)
    if np.iscomplexobj(arr):
        raise NotImplementedError('complex tensor trains are not supported')
    return torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)).to(dev)


def from_host_batch(core_lists):
    """Many tensor trains from host cores with ONE host-to-device copy.

    Every core of every train is packed into one pinned host buffer, copied
    once, and the device cores are views into that single allocation (a
    batch of small trains would otherwise cost one copy per core).
    """
    dev = device()
    shapes = [[tuple(np.shape(c)) for c in cl] for cl in core_lists]
    for sh in shapes:
        if not sh or not all(len(s) == 4 for s in sh):
            raise ValueError('List elements must be 4-dimensional arrays.')
        if not all(sh[i][3] == sh[i + 1][0] for i in range(len(sh) - 1)):
            raise ValueError('Shapes of list elements do not match.')
    total = sum(int(np.prod(s)) for sh in shapes for s in sh)
    host = torch.empty(total, dtype=torch.float64, pin_memory=True)
    hv = host.numpy()
    off = 0
    for cl in core_lists:
        for c in cl:
            a = np.asarray(c)
            if np.iscomplexobj(a):
                raise NotImplementedError('complex tensor trains are not supported')
            hv[off:off + a.size] = a.reshape(-1)
            off += a.size
    buf = host.to(dev, non_blocking=True)
    torch.cuda.current_stream(dev).synchronize()  # the pinned source must outlive the copy
    # trains as lazy views: per-core tensors are made only when a core is used
    out, off = [], 0
    for sh in shapes:
        sizes = np.array([int(np.prod(s)) for s in sh], dtype=np.int64)
        offs = off + np.concatenate(([0], np.cumsum(sizes)[:-1]))
        off += int(sizes.sum())
        out.append(TT._from_flat(buf, offs, [s[1] for s in sh], [s[2] for s in sh],
                                 [s[0] for s in sh] + [sh[-1][3]]))
    return out


def _host_readonly(c):
    a = c.detach().cpu().numpy()
    a.setflags(write=False)
    return a


class _CoreList:
    """List-like view of the cores: reads give READ-ONLY host numpy copies
    (an in-place edit such as ``t.cores[i][...] = v`` raises instead of being
    lost on a copy); assigning ``t.cores[i] = arr`` uploads."""

    def __init__(self, tt):
        self._tt = tt

    def __len__(self):
        return self._tt.order

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [_host_readonly(c) for c in self._tt.device_cores[i]]
        return _host_readonly(self._tt.device_cores[i])

    def __setitem__(self, i, value):
        t = self._tt
        c = _as_device_core(value, t.device_cores[i].device)
        if c.dim() != 4:
            raise ValueError('List elements must be 4-dimensional arrays.')
        t.device_cores[i] = c
        t.row_dims[i], t.col_dims[i] = int(c.shape[1]), int(c.shape[2])
        t.ranks[i], t.ranks[i + 1] = int(c.shape[0]), int(c.shape[3])

    def __iter__(self):
        for i in range(len(self)):
            yield self[i]


class TT:
    """Tensor train / TT operator with cores on the HIP device.

    Constructor semantics follow scikit_tt/tensor_train.py:149-267: for a
    list of cores every element must be 4-D (ValueError) and neighbouring
    ranks must match (ValueError); a full array of shape (m_1..m_d,
    n_1..n_d) is decomposed by TT-SVD on the device (truncated SVDs with
    ``threshold`` / ``max_rank``); anything else raises TypeError.
    """

    def __init__(self, x, threshold=0, max_rank=np.inf, progress=False, string=None):
        if isinstance(x, list):
            shapes = [tuple(c.shape) for c in x]
            if not all(len(s) == 4 for s in shapes):
                raise ValueError('List elements must be 4-dimensional arrays.')
            if not all(shapes[i][3] == shapes[i + 1][0] for i in range(len(shapes) - 1)):
                raise ValueError('Shapes of list elements do not match.')
            dev = device()
            self.device_cores = [_as_device_core(c, dev) for c in x]
            self.order = len(x)
            self.row_dims = [s[1] for s in shapes]
            self.col_dims = [s[2] for s in shapes]
            self.ranks = [s[0] for s in shapes] + [shapes[-1][3]]
            if threshold != 0 or max_rank != np.inf:
                self.ortho(threshold=threshold, max_rank=max_rank)
        elif isinstance(x, np.ndarray):
            t = _tt_svd(x, threshold, max_rank)
            self.__dict__.update(t.__dict__)
        else:
            raise TypeError('Parameter must be either a list of cores or an ndarray.')

    # -- attributes -----------------------------------------------------------

    @property
    def cores(self):
        return _CoreList(self)

    @cores.setter
    def cores(self, value):
        if not isinstance(value, list):
            raise TypeError('cores must be a list')
        t = TT(value)
        self.__dict__.update(t.__dict__)

    @property
    def device(self):
        if self.__dict__.get('_cores') is None and self.__dict__.get('_lazy') is not None:
            return self._lazy[0].device  # without creating the per-core tensors
        return self.device_cores[0].device

    def __repr__(self):
        return ('\n'
                'Tensor train with order    = {d}, \n'
                '                  row_dims = {m}, \n'
                '                  col_dims = {n}, \n'
                '                  ranks    = {r}'.format(d=self.order, m=self.row_dims, n=self.col_dims,
                                                          r=self.ranks))

    def isoperator(self):
        """tensor_train.py:933-946"""
        return not (all(i == 1 for i in self.row_dims) or all(i == 1 for i in self.col_dims))

    def _desc(self):
        return _lib.DescPack(self.order, self.row_dims, self.col_dims, self.ranks, self._core_ptrs())

    @classmethod
    def _from_device(cls, cores):
        t = cls.__new__(cls)
        t.device_cores = cores
        t.order = len(cores)
        t.row_dims = [int(c.shape[1]) for c in cores]
        t.col_dims = [int(c.shape[2]) for c in cores]
        t.ranks = [int(c.shape[0]) for c in cores] + [int(cores[-1].shape[3])]
        return t

    @classmethod
    def _from_flat(cls, buf, offsets, row_dims, col_dims, ranks):
        """Train whose core i is buf[offsets[i]:...] viewed as (r_i, m_i, n_i, r_i+1).

        The per-core tensors are created on first access, so that a batch of
        results costs no per-core work until a result is used.
        """
        t = cls.__new__(cls)
        t._cores = None
        t._lazy = (buf, offsets)
        t.order = len(row_dims)
        t.row_dims, t.col_dims, t.ranks = list(row_dims), list(col_dims), list(ranks)
        return t

    @property
    def device_cores(self):
        cores = self.__dict__.get('_cores')
        if cores is None:
            buf, offs = self._lazy
            r, m, n = self.ranks, self.row_dims, self.col_dims
            cores = []
            for i in range(self.order):
                shape = (r[i], m[i], n[i], r[i + 1])
                o = int(offs[i])
                cores.append(buf[o:o + r[i] * m[i] * n[i] * r[i + 1]].view(shape))
            self._cores = cores
            self._lazy = None
        return cores

    @device_cores.setter
    def device_cores(self, cores):
        self._cores = cores
        self._lazy = None

    def _core_ptrs(self):
        """Device addresses of the cores (without creating per-core tensors)."""
        return self._core_ptr_array().tolist()

    def _core_ptr_array(self):
        """Device addresses of the cores as a uint64 array."""
        if self.__dict__.get('_cores') is None and self.__dict__.get('_lazy') is not None:
            buf, offs = self._lazy
            return np.uint64(buf.data_ptr()) + np.uint64(8) * np.asarray(offs, dtype=np.uint64)
        return np.array([c.data_ptr() for c in self.device_cores], dtype=np.uint64)

    # -- copies and conversion ---------------------------------------------------

    def copy(self):
        """Deep copy (tensor_train.py:948-964)."""
        return TT._from_device([c.clone() for c in self.device_cores])

    def full(self):
        """Full tensor m_1..m_d x n_1..n_d as a host array (tensor_train.py:1029-1059)."""
        if self.ranks[0] != 1 or self.ranks[-1] != 1:
            raise ValueError('The first and last rank have to be 1!')
        m = self.matricize()
        d = self.order
        t = m.reshape(list(self.row_dims) + list(self.col_dims))
        return t

    def matricize(self):
        """Matrix (operator) or vector (tensor) on the host (tensor_train.py:1061-1090).

        Contracted on the device (plain tensordot of the cores: a format
        conversion, not part of the ALS path), then copied to the host.
        """
        c0 = self.device_cores[0]
        t = c0.reshape(c0.shape[1], c0.shape[2], c0.shape[3])
        for c in self.device_cores[1:]:
            u = torch.tensordot(t, c, dims=([2], [0]))  # M N m n r'
            u = u.permute(0, 2, 1, 3, 4)
            t = u.reshape(u.shape[0] * u.shape[1], u.shape[2] * u.shape[3], u.shape[4])
        m, n = int(np.prod(self.row_dims)), int(np.prod(self.col_dims))
        out = t.reshape(m) if n == 1 else t.reshape(m, n)
        return out.cpu().numpy()

    # -- arithmetic (tensor_train.py:282-524) --------------------------------------

    def __add__(self, other):
        """Block-diagonal rank concatenation (tensor_train.py:282-344)."""
        if not isinstance(other, TT):
            raise TypeError('Unsupported parameter.')
        if self.row_dims != other.row_dims or self.col_dims != other.col_dims:
            raise ValueError('Tensor trains must have the same dimensions')
        return _embed_sum(self, other, 1.0)

    def __sub__(self, other):
        """self + (-1) * other (tensor_train.py:346-366)."""
        if not isinstance(other, TT):
            raise TypeError('Unsupported parameter.')
        if self.row_dims != other.row_dims or self.col_dims != other.col_dims:
            raise ValueError('Tensor trains must have the same dimensions')
        return _embed_sum(self, other, -1.0)

    def __mul__(self, scalar):
        """Scalar multiple: core 0 is scaled (tensor_train.py:368-400)."""
        if isinstance(scalar, bool) or not isinstance(scalar, (numbers.Real,)):
            if isinstance(scalar, complex):
                raise NotImplementedError('complex scalars are not supported')
            raise TypeError('Unsupported parameter.')
        out = self.copy()
        c = out.device_cores[0]
        r0, m, n, r1 = c.shape
        dst = torch.empty_like(c)
        _lib.check(_lib.load().ttals_core_embed(_stream(c.device), r0, m * n, r1, ctypes.c_void_p(c.data_ptr()),
                                                 float(scalar), r0, r1, 0, 0, ctypes.c_void_p(dst.data_ptr())),
                   what='core_embed')
        out.device_cores[0] = dst
        return out

    def __rmul__(self, scalar):
        """tensor_train.py:402-420"""
        return self.__mul__(scalar)

    def __matmul__(self, other):
        """TT product, core by core (tensor_train.py:422-503)."""
        if not isinstance(other, TT):
            raise TypeError('Unsupported argument.')
        if self.col_dims != other.row_dims:
            raise ValueError('Dimensions do not match.')
        lib = _lib.load()
        cores = []
        for c1, c2 in zip(self.device_cores, other.device_cores):
            r1, m, n, r2 = c1.shape
            s1, _, p, s2 = c2.shape
            out = torch.empty((r1 * s1, m, p, r2 * s2), dtype=torch.float64, device=c1.device)
            _lib.check(lib.ttals_core_product(_stream(c1.device), r1, m, n, r2, s1, p, s2,
                                              ctypes.c_void_p(c1.data_ptr()), ctypes.c_void_p(c2.data_ptr()),
                                              ctypes.c_void_p(out.data_ptr())), what='core_product')
            cores.append(out)
        prod = TT._from_device(cores)
        if np.prod(prod.row_dims) == 1 and np.prod(prod.col_dims) == 1:
            v = cores[0].reshape(cores[0].shape[0], -1)
            for c in cores[1:]:
                v = v @ c.reshape(c.shape[0], -1)
            return float(v[0, 0].item())
        return prod

    def dot(self, other):
        """Alias of ``@`` (tensor_train.py:505-524)."""
        return self.__matmul__(other)

    # -- orthonormalisation and norms (tensor_train.py:1092-1430) ----------------

    def ortho_right(self, start_index=None, end_index=1, threshold=0, max_rank=np.inf):
        """Right-orthonormalisation (tensor_train.py:1207-1310), in place.

        Without truncation the device runs a Householder-QR sweep: the
        reference keeps all min(r_i, m_i n_i r_{i+1}) thin-SVD triplets, i.e.
        the same ranks and tensor (only the orthonormal gauge differs).  With
        ``threshold > 0`` or a finite ``max_rank`` it adds a Jacobi SVD of
        the R factor and keeps s/s0 > threshold, then at most max_rank.
        """
        if start_index is None:
            start_index = self.order - 1
        caps = _check_ortho_args(self.order, start_index, end_index, threshold, max_rank)
        if start_index >= end_index:
            self._ortho_sweep(1, int(start_index), int(end_index), float(threshold), caps)
        return self

    def ortho_left(self, start_index=0, end_index=None, threshold=0.0, max_rank=np.inf, progress=False,
                   string='Left-orthonormalization'):
        """Left-orthonormalisation (tensor_train.py:1092-1205), in place."""
        if end_index is None:
            end_index = self.order - 2
        caps = _check_ortho_args(self.order, start_index, end_index, threshold, max_rank)
        if end_index >= start_index:
            self._ortho_sweep(0, int(start_index), int(end_index), float(threshold), caps)
        return self

    def _ortho_sweep(self, direction, start, end, threshold, caps):
        ctx = _lib.context(self.device.index)
        pack = self._desc()
        cap_arr = None if caps is None else (ctypes.c_int * (self.order + 1))(*caps)
        _lib.check(_lib.load().ttals_tt_ortho(ctx, _stream(self.device), ctypes.byref(pack.desc), direction, start,
                                              end, threshold, cap_arr), ctx, 'tt_ortho')
        ranks = pack.final_ranks()
        for i in range(self.order):
            c = self.device_cores[i]
            shape = (ranks[i], self.row_dims[i], self.col_dims[i], ranks[i + 1])
            if tuple(c.shape) != shape:
                self.device_cores[i] = c.reshape(-1)[:int(np.prod(shape))].reshape(shape).clone()
        self.ranks = ranks

    def ortho(self, threshold=0, max_rank=np.inf):
        """ortho_left then ortho_right (tensor_train.py:1312-1332)."""
        return self.ortho_left(threshold=threshold, max_rank=np.inf).ortho_right(threshold=threshold,
                                                                               max_rank=max_rank)

    def norm(self, p=2):
        """Norm (tensor_train.py:1334-1430): p=2 the Euclidean/Frobenius norm
        (device QR sweep), p=1 the Manhattan norm / maximum column sum of a
        non-negative train (device core products)."""
        if p == 2:
            ctx = _lib.context(self.device.index)
            pack = self._desc()
            out = ctypes.c_double()
            _lib.check(_lib.load().ttals_tt_norm(ctx, _stream(self.device), ctypes.byref(pack.desc),
                                                 ctypes.byref(out)), ctx, 'tt_norm')
            return float(out.value)
        if p == 1:
            # Manhattan norm (tensor) / maximum column sum (operator), entries
            # assumed non-negative as in the reference: the row modes are
            # summed by a core product with all-ones row cores (1, 1, m, 1)
            t = self.transpose() if all(m == 1 for m in self.row_dims) else self
            dev = t.device
            ones = TT._from_device([torch.ones((1, 1, m, 1), dtype=torch.float64, device=dev) for m in t.row_dims])
            s = ones.dot(t)
            if isinstance(s, float):
                return s
            return float(np.max(s.matricize()))
        raise ValueError('p must be 1 or 2.')

    def transpose(self):
        """Swap row and column modes of every core (tensor_train.py:795-860, conjugate=False)."""
        return TT._from_device([c.permute(0, 2, 1, 3).contiguous() for c in self.device_cores])


    def tt2qtt(self, row_dims, col_dims, threshold=0):
        """Split every core into a chain of smaller cores (QTT format,
        tensor_train.py:1432-1516): core i with row/col dims m_i = prod
        row_dims[i], n_i = prod col_dims[i] is split left to right by
        truncated SVDs on the device (s/s0 > threshold when threshold != 0)."""
        from .utils import truncated_svd
        qtt = []
        for i in range(self.order):
            core = self.device_cores[i]
            rank = self.ranks[i]
            row_dim, col_dim = self.row_dims[i], self.col_dims[i]
            r_right = self.ranks[i + 1]
            for j in range(len(row_dims[i]) - 1):
                row_dim //= row_dims[i][j]
                col_dim //= col_dims[i][j]
                c6 = core.reshape(rank, row_dims[i][j], row_dim, col_dims[i][j], col_dim, r_right)
                c6 = c6.permute(0, 1, 3, 2, 4, 5).contiguous()
                u, s, v = truncated_svd(c6.reshape(rank * row_dims[i][j] * col_dims[i][j], -1), threshold=threshold)
                k = s.shape[0]
                qtt.append(u.reshape(rank, row_dims[i][j], col_dims[i][j], k))
                core = s[:, None] * v
                rank = k
            qtt.append(core.reshape(rank, row_dim, col_dim, r_right).contiguous())
        return TT._from_device(qtt)


# ---------------------------------------------------------------------------
# helpers


def _tt_svd(x, threshold, max_rank):
    """TT decomposition of a full array (tensor_train.py:200-261): modes
    interleaved (m_i, n_i), then split left to right by truncated SVDs on the
    device; singular values pushed right."""
    from .utils import truncated_svd
    if x.ndim % 2:
        raise ValueError('Number of dimensions must be a multiple of 2.')
    if np.iscomplexobj(x):
        raise NotImplementedError('complex tensor trains are not supported')
    order = x.ndim // 2
    row_dims, col_dims = x.shape[:order], x.shape[order:]
    perm = [order * j + i for i in range(order) for j in range(2)]
    y = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(device()).permute(perm).contiguous()
    ranks = [1] * (order + 1)
    cores = []
    for i in range(order - 1):
        y = y.reshape(ranks[i] * row_dims[i] * col_dims[i], -1)
        u, s, v = truncated_svd(y, threshold=threshold, max_rank=max_rank)
        ranks[i + 1] = s.shape[0]
        cores.append(u.reshape(ranks[i], row_dims[i], col_dims[i], ranks[i + 1]).contiguous())
        y = s[:, None] * v
    cores.append(y.reshape(ranks[-2], row_dims[-1], col_dims[-1], 1).contiguous())
    return TT._from_device(cores)


def _check_ortho_args(order, start_index, end_index, threshold, max_rank):
    """Argument checks of ortho_left/right (tensor_train.py:1137-1159); returns the
    per-bond rank caps for the device (0 = unbounded) or None."""
    if not isinstance(start_index, (int, np.integer)) or not isinstance(end_index, (int, np.integer)):
        raise TypeError('Start and end indices must be integers.')
    if not isinstance(threshold, (int, float, np.floating, np.integer)) or threshold < 0:
        raise ValueError('Threshold must be greater or equal 0.')
    if (isinstance(max_rank, (int, np.integer)) and max_rank > 0) or max_rank == np.inf:
        caps = [1] + [max_rank] * (order - 1) + [1]
    elif isinstance(max_rank, list) and len(max_rank) == order + 1 and all(
            (isinstance(r, (int, np.integer)) and r > 0) or r == np.inf for r in max_rank):
        caps = list(max_rank)
    else:
        raise ValueError('Maximum rank(s) must be positive integers.')
    if all(c == np.inf for c in caps[1:-1]):
        return None
    return [0 if c == np.inf else int(c) for c in caps]


def _embed_sum(x, y, sign):
    """x + sign*y with the rank layout of tensor_train.py:313-333."""
    lib = _lib.load()
    d = x.order
    rs = [1] + [x.ranks[i] + y.ranks[i] for i in range(1, d)] + [1]
    cores = []
    for i in range(d):
        cx, cy = x.device_cores[i], y.device_cores[i]
        m, n = cx.shape[1], cx.shape[2]
        dst = torch.zeros((rs[i], m, n, rs[i + 1]), dtype=torch.float64, device=cx.device)
        st = _stream(cx.device)
        _lib.check(lib.ttals_core_embed(st, cx.shape[0], m * n, cx.shape[3], ctypes.c_void_p(cx.data_ptr()), 1.0,
                                        rs[i], rs[i + 1], 0, 0, ctypes.c_void_p(dst.data_ptr())), what='embed')
        s = sign if i == 0 else 1.0
        _lib.check(lib.ttals_core_embed(st, cy.shape[0], m * n, cy.shape[3], ctypes.c_void_p(cy.data_ptr()), s,
                                        rs[i], rs[i + 1], rs[i] - cy.shape[0], rs[i + 1] - cy.shape[3],
                                        ctypes.c_void_p(dst.data_ptr())), what='embed')
        cores.append(dst)
    return TT._from_device(cores)


def _rank_list(order, ranks):
    if not isinstance(ranks, list):
        ranks = [1] + [ranks] * (order - 1) + [1]
    return ranks


def zeros(row_dims, col_dims, ranks=1):
    """tensor_train.py:1806-1835"""
    r = _rank_list(len(row_dims), ranks)
    dev = device()
    return TT._from_device([torch.zeros((r[i], row_dims[i], col_dims[i], r[i + 1]), dtype=torch.float64,
                                        device=dev) for i in range(len(row_dims))])


def ones(row_dims, col_dims, ranks=1):
    """tensor_train.py:1838-1867"""
    r = _rank_list(len(row_dims), ranks)
    dev = device()
    return TT._from_device([torch.ones((r[i], row_dims[i], col_dims[i], r[i + 1]), dtype=torch.float64,
                                       device=dev) for i in range(len(row_dims))])


def eye(dims):
    """tensor_train.py:1870-1893"""
    dev = device()
    return TT._from_device([torch.eye(n, dtype=torch.float64, device=dev).reshape(1, n, n, 1) for n in dims])


def rand(row_dims, col_dims, ranks=1):
    """Uniform [0,1) cores from numpy's global generator (tensor_train.py:1921-1950)."""
    r = _rank_list(len(row_dims), ranks)
    return TT([np.random.rand(r[i], row_dims[i], col_dims[i], r[i + 1]) for i in range(len(row_dims))])


def residual_error(operator, lhs, rhs):
    """||operator @ lhs - rhs|| (tensor_train.py:2035-2074).

    Same quantity as the reference's core-wise SVD compression; computed on
    the device by one right-to-left Householder sweep over the block cores
    of operator @ lhs - rhs, generated on the fly (never stored).
    """
    return float(residual_norms([operator], [lhs], [rhs])[0][0])


def residual_norms(operators, lhss, rhss):
    """Residual norms of many systems in ONE launch.

    Returns two float64 arrays: ``||A_s x_s - b_s||`` and ``||b_s||``, the
    numerator and denominator of the reference's rel.res check
    ``(A.dot(x) - b).norm() / b.norm()`` (tests/test_sle.py:52-53).
    """
    count = len(operators)
    if count == 0 or len(lhss) != count or len(rhss) != count:
        raise ValueError('operators, solutions and right-hand sides must be non-empty lists of equal length')
    for t in list(operators) + list(lhss) + list(rhss):
        if not isinstance(t, TT):
            raise TypeError('Unsupported argument.')
    dev = operators[0].device
    dA = _lib.DescBatch.from_tts(operators)
    dX = _lib.DescBatch.from_tts(lhss)
    dB = _lib.DescBatch.from_tts(rhss)
    res = np.zeros(count)
    bn = np.zeros(count)
    ctx = _lib.context(dev.index)
    _lib.check(_lib.load().ttals_residual_norm_batch(
        ctx, _stream(dev), count, dA.ptr, dX.ptr, dB.ptr, res.ctypes.data_as(_lib.c_double_p),
        bn.ctypes.data_as(_lib.c_double_p)), ctx, 'residual_norm_batch')
    return res, bn
