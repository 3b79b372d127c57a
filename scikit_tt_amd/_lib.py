"""ctypes binding of libttals.so (the C ABI declared in include/ttals.h).

The shared library is built in-tree (``python -m scikit_tt_amd.build``) and
loaded from this package directory.  There is no fallback: if the library or
a HIP device is missing, every compute call raises.
"""
import contextlib
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libttals.so')

E_ARG, E_SHAPE, E_HIP, E_OOM, E_UNSUPPORTED = -1, -2, -3, -4, -5
S_SINGULAR, S_NOCONV, S_UNSUPPORTED, S_NONFINITE, S_NONSYM, S_DEADLINE = 1, 2, 3, 4, 5, 7

c_int_p = ctypes.POINTER(ctypes.c_int)
c_double_p = ctypes.POINTER(ctypes.c_double)


class TTDesc(ctypes.Structure):
    _fields_ = [('order', ctypes.c_int),
                ('row_dims', c_int_p),
                ('col_dims', c_int_p),
                ('ranks', c_int_p),
                ('cores', ctypes.POINTER(ctypes.c_void_p))]


class Result(ctypes.Structure):
    _fields_ = [('status', ctypes.c_int),
                ('sweeps', ctypes.c_int),
                ('unconverged', ctypes.c_int),
                ('max_iters', ctypes.c_int),
                ('krylov_iters', ctypes.c_longlong),
                ('local_solves', ctypes.c_longlong),
                ('lambda_', ctypes.c_double),
                ('worst_residual', ctypes.c_double),
                ('flops', ctypes.c_double),
                ('cycles', ctypes.c_double * 7),
                ('device_ms', ctypes.c_double),
                ('bytes', ctypes.c_double)]


# exported symbol -> (restype, argtypes)
_vp, _i, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_TTp = ctypes.POINTER(TTDesc)
_Rp = ctypes.POINTER(Result)
SIGNATURES = {
    'ttals_version': (_i, []),
    'ttals_ctx_create': (_i, [_i, ctypes.POINTER(_vp)]),
    'ttals_ctx_destroy': (None, [_vp]),
    'ttals_last_error': (ctypes.c_char_p, [_vp]),
    'ttals_sle_als_batch': (_i, [_vp, _vp, _i, _TTp, _TTp, _TTp, _TTp, _i, _i, _d, _i, _Rp]),
    'ttals_evp_als_batch': (_i, [_vp, _vp, _i, _TTp, _TTp, _TTp, _i, _d, _i, _d, _d, _i, _Rp]),
    'ttals_sle_als_gang': (_i, [_vp, _vp, _i, _TTp, _TTp, _TTp, _TTp, _i, _i, _d, _i, _i, _i, _Rp]),
    'ttals_sle_als_batch_static': (_i, [_vp, _vp, _i, _TTp, _TTp, _TTp, _TTp, _i, _i, _d, _i, _Rp]),
    'ttals_sle_submit': (_i, [_vp, _vp, _i, _TTp, _TTp, _TTp, _TTp, _i, _i, _d, _i]),
    'ttals_evp_submit': (_i, [_vp, _vp, _i, _TTp, _TTp, _TTp, _i, _d, _i, _d, _d, _i]),
    'ttals_query': (_i, [_vp]),
    'ttals_wait': (_i, [_vp, _Rp]),
    'ttals_last_launch_ms': (_i, [_vp, ctypes.POINTER(ctypes.c_double)]),
    'ttals_set_arena_kb': (_i, [_vp, _i]),
    'ttals_set_local_tol_policy': (_i, [_vp, _i]),
    'ttals_set_matvec_variant': (_i, [_vp, _i]),
    'ttals_set_schedule': (_i, [_vp, _i, _i]),
    'ttals_dense_eig': (_i, [_vp, _vp, _i, _vp, _vp, _i, _i, _d, _i, _i, _vp, _vp]),
    'ttals_set_time_budget': (_i, [_vp, _d]),
    'ttals_set_cost_model': (_i, [_vp, _vp, _i]),
    'ttals_get_cost_model': (_i, [_vp, _vp, _i]),
    'ttals_sle_cost': (_i, [_vp, _i, _TTp, _TTp, _vp]),
    'ttals_last_plan': (_i, [_vp, _vp, _i]),
    'ttals_last_plan_ms': (_i, [_vp, ctypes.POINTER(ctypes.c_double)]),
    'ttals_sle_plan': (_i, [_vp, _i, _TTp, _TTp, _i, _i, _vp, _vp]),
    'ttals_stack_left_op': (_i, [_vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    'ttals_stack_right_op': (_i, [_vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    'ttals_stack_left_rhs': (_i, [_vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    'ttals_stack_right_rhs': (_i, [_vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    'ttals_micro_matrix': (_i, [_vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    'ttals_micro_matvec': (_i, [_vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    'ttals_micro_rhs': (_i, [_vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    'ttals_local_solve': (_i, [_vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _d, _i, c_int_p]),
    'ttals_local_eig': (_i, [_vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _d, _d, _i, c_double_p,
                             c_int_p]),
    'ttals_qr': (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp]),
    'ttals_core_product': (_i, [_vp, _i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp]),
    'ttals_core_embed': (_i, [_vp, _i, _i, _i, _vp, _d, _i, _i, _i, _i, _vp]),
    'ttals_tt_norm': (_i, [_vp, _vp, _TTp, c_double_p]),
    'ttals_tt_ortho': (_i, [_vp, _vp, _TTp, _i, _i, _i, _d, c_int_p]),
    'ttals_set_trace': (_i, [_vp, _vp, ctypes.c_longlong]),
    'ttals_residual_norm_batch': (_i, [_vp, _vp, _i, _TTp, _TTp, _TTp, c_double_p, c_double_p]),
    'ttals_truncated_svd': (_i, [_vp, _vp, _i, _i, _vp, _d, _i, _i, _vp, _vp, _vp, c_int_p]),
    'ttals_core_merge': (_i, [_vp, _i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp]),
}

_lock = threading.Lock()
_lib = None
_ctx = {}


def load(path=LIB_PATH):
    """Load libttals.so and declare every exported signature."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise ImportError(
                    'libttals.so not found at %s: build it with `python -m scikit_tt_amd.build`' % path)
            lib = ctypes.CDLL(path)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def context(device):
    """Per-device engine context (workspace owner), created on first use."""
    lib = load()
    with _lock:
        c = _ctx.get(device)
        if c is None:
            h = ctypes.c_void_p()
            rc = lib.ttals_ctx_create(int(device), ctypes.byref(h))
            if rc != 0:
                raise RuntimeError('ttals_ctx_create(%d) failed with %d' % (device, rc))
            c = h
            _ctx[device] = c
    return c


@contextlib.contextmanager
def schedule(ctx, mode, gang_precision=0):
    """The context's sle batch schedule set to ``mode`` for the block, then
    back to one workgroup per system (the contexts are shared per device)."""
    lib = load()
    check(lib.ttals_set_schedule(ctx, int(mode), int(gang_precision)), ctx, 'set_schedule')
    try:
        yield
    finally:
        lib.ttals_set_schedule(ctx, 0, 0)


def new_context(device):
    """A dedicated engine context (own workspace; one pending submission at a time)."""
    lib = load()
    h = ctypes.c_void_p()
    rc = lib.ttals_ctx_create(int(device), ctypes.byref(h))
    if rc != 0:
        raise RuntimeError('ttals_ctx_create(%d) failed with %d' % (device, rc))
    return h


def check(rc, ctx=None, what='ttals'):
    """Map a negative C-ABI return code to the Python exception the reference raises."""
    if rc >= 0:
        return rc
    msg = what
    if ctx is not None:
        err = load().ttals_last_error(ctx)
        if err:
            msg = '%s: %s' % (what, err.decode())
    if rc in (E_ARG, E_SHAPE):
        raise ValueError(msg)
    if rc == E_OOM:
        raise MemoryError(msg)
    if rc == E_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError('%s (code %d)' % (msg, rc))


# ttals_tt as a numpy record (same layout as the C struct: int, pad, 4 pointers)
TT_DTYPE = np.dtype([('order', '<i4'), ('pad', '<i4'), ('row_dims', '<u8'), ('col_dims', '<u8'),
                     ('ranks', '<u8'), ('cores', '<u8')])
assert TT_DTYPE.itemsize == ctypes.sizeof(TTDesc)
assert [TT_DTYPE.fields[f][1] for f in ('order', 'row_dims', 'col_dims', 'ranks', 'cores')] == \
    [getattr(TTDesc, f).offset for f in ('order', 'row_dims', 'col_dims', 'ranks', 'cores')]


class DescBatch:
    """ttals_tt[count] for many trains, built in bulk.

    One byte buffer holds the record array (same layout as the C struct)
    followed by every train's dims, ranks and core pointers; ``self.c`` is a
    ctypes ``ttals_tt`` array created over that buffer, so whatever holds
    ``self.c`` (or this object) keeps all of it alive.  ``ranks_of(s)`` reads
    train s's ranks (the library writes final ranks of outputs there).
    """

    def __init__(self, orders, rows, cols, ranks, ptrs):
        orders = np.asarray(orders, dtype=np.int64)
        n = len(orders)
        nd, nr = int(orders.sum()), int((orders + 1).sum())
        rec = n * TT_DTYPE.itemsize
        o_rows = rec
        o_cols = o_rows + 4 * nd
        o_ranks = o_cols + 4 * nd
        o_ptrs = (o_ranks + 4 * nr + 7) // 8 * 8
        self.blob = np.zeros(o_ptrs + 8 * nd, dtype=np.uint8)
        arr = self.blob[:rec].view(TT_DTYPE)
        self.rows = self.blob[o_rows:o_cols].view(np.int32)
        self.cols = self.blob[o_cols:o_ranks].view(np.int32)
        self.ranks = self.blob[o_ranks:o_ranks + 4 * nr].view(np.int32)
        self.ptrs = self.blob[o_ptrs:].view(np.uint64)
        self.rows[:] = rows
        self.cols[:] = cols
        self.ranks[:] = ranks
        self.ptrs[:] = ptrs
        self.orders = orders
        self.doff = np.concatenate(([0], np.cumsum(orders)[:-1])).astype(np.int64)
        self.roff = np.concatenate(([0], np.cumsum(orders + 1)[:-1])).astype(np.int64)
        base = self.blob.ctypes.data
        arr['order'] = orders
        arr['row_dims'] = base + o_rows + 4 * self.doff
        arr['col_dims'] = base + o_cols + 4 * self.doff
        arr['ranks'] = base + o_ranks + 4 * self.roff
        arr['cores'] = base + o_ptrs + 8 * self.doff
        self.c = (TTDesc * n).from_buffer(self.blob)

    @property
    def ptr(self):
        return self.c

    @classmethod
    def from_tts(cls, tts):
        lazy = [t.__dict__.get('_lazy') if t.__dict__.get('_cores') is None else None for t in tts]
        if lazy[0] is not None and all(z is not None and z[0] is lazy[0][0] for z in lazy):
            # views of one flat buffer (e.g. from_host_batch): one base address
            offs = np.concatenate([z[1] for z in lazy]).astype(np.uint64)
            ptrs = np.uint64(lazy[0][0].data_ptr()) + np.uint64(8) * offs
        else:
            ptrs = np.concatenate([t._core_ptr_array() for t in tts])
        t0 = tts[0]
        if all(t.row_dims == t0.row_dims and t.col_dims == t0.col_dims and t.ranks == t0.ranks for t in tts):
            n = len(tts)  # one structure (the usual batch): dims tiled
            return cls([t0.order] * n, np.tile(t0.row_dims, n), np.tile(t0.col_dims, n), np.tile(t0.ranks, n), ptrs)
        orders, rows, cols, ranks = [], [], [], []
        for t in tts:
            orders.append(t.order)
            rows += t.row_dims
            cols += t.col_dims
            ranks += t.ranks
        return cls(orders, rows, cols, ranks, ptrs)

    def ranks_of(self, s):
        o = int(self.roff[s])
        return self.ranks[o:o + int(self.orders[s]) + 1].tolist()


class DescPack:
    """Keeps the ctypes arrays behind a TTDesc alive."""

    def __init__(self, order, row_dims, col_dims, ranks, ptrs):
        self.rows = (ctypes.c_int * order)(*row_dims)
        self.cols = (ctypes.c_int * order)(*col_dims)
        self.ranks = (ctypes.c_int * (order + 1))(*ranks)
        self.ptrs = (ctypes.c_void_p * order)(*ptrs)
        self.desc = TTDesc(order, self.rows, self.cols, self.ranks, self.ptrs)

    def final_ranks(self):
        return [int(r) for r in np.ctypeslib.as_array(self.ranks)]
