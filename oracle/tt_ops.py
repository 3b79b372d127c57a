"""Oracle: tensor-train algebra on plain lists of numpy cores (TEST ONLY).

A TT is a Python list of 4-D float64 arrays ``core[i]`` of shape
``(r_i, m_i, n_i, r_{i+1})`` (scikit_tt/tensor_train.py:14-21).  Functions
return new lists; the ortho* functions also return the (possibly changed)
rank list so callers can mirror the reference's in-place ``ranks`` updates.
"""
import numpy as np
import scipy.linalg as sla


def validate(cores):
    """Shape checks of the list constructor (tensor_train.py:175-198)."""
    if not isinstance(cores, list):
        raise TypeError('Parameter must be either a list of cores or an ndarray.')
    if not all(c.ndim == 4 for c in cores):
        raise ValueError('List elements must be 4-dimensional arrays.')
    for left, right in zip(cores[:-1], cores[1:]):
        if left.shape[3] != right.shape[0]:
            raise ValueError('Shapes of list elements do not match.')
    return cores


def ranks_of(cores):
    """Rank list r_0..r_d (tensor_train.py:187)."""
    return [c.shape[0] for c in cores] + [cores[-1].shape[3]]


def copy(cores):
    """Deep copy (tensor_train.py:948-964)."""
    return [np.array(c, copy=True) for c in cores]


def add(x, y):
    """Block-diagonal rank concatenation (tensor_train.py:282-344).

    Interior cores place ``x`` in the upper-left and ``y`` in the lower-right
    rank block; the outer ranks are forced to 1, so core 0 is a row block and
    the last core a column block.
    """
    d = len(x)
    rx, ry = ranks_of(x), ranks_of(y)
    if [c.shape[1] for c in x] != [c.shape[1] for c in y] or \
            [c.shape[2] for c in x] != [c.shape[2] for c in y]:
        raise ValueError('Tensor trains must have the same dimensions')
    rs = [1] + [rx[i] + ry[i] for i in range(1, d)] + [1]
    out = []
    for i in range(d):
        dt = np.result_type(x[i], y[i], np.float64)
        c = np.zeros((rs[i], x[i].shape[1], x[i].shape[2], rs[i + 1]), dtype=dt)
        c[:rx[i], :, :, :rx[i + 1]] = x[i]
        c[rs[i] - ry[i]:, :, :, rs[i + 1] - ry[i + 1]:] = y[i]
        out.append(c)
    return out


def scale(x, s):
    """Scalar multiple; only core 0 is scaled (tensor_train.py:368-420)."""
    if not isinstance(s, (int, float, complex, np.floating)):
        raise TypeError('Unsupported parameter.')
    out = copy(x)
    out[0] = s * out[0]
    return out


def sub(x, y):
    """x - y = x + (-1)*y (tensor_train.py:346-366)."""
    return add(x, scale(y, -1))


def core_product(c1, c2):
    """Product of an operator core and a core (tensor_train.py:447-481).

    ``(r1, m, n, r2) x (s1, n, p, s2) -> (r1*s1, m, p, r2*s2)`` with the
    product rank index ``i_self * s + i_other``.
    """
    r1, m, n, r2 = c1.shape
    s1, n2, p, s2 = c2.shape
    if n != n2:
        raise ValueError('Dimensions do not match.')
    # out[i, k, m, p, j, l] = sum_n c1[i, m, n, j] c2[k, n, p, l]
    t = np.einsum('imnj,knpl->ikmpjl', c1, c2, optimize=True)
    return t.reshape(r1 * s1, m, p, r2 * s2)


def dot(x, y):
    """TT product ``x @ y`` (tensor_train.py:422-503); scalar if all dims 1."""
    if [c.shape[2] for c in x] != [c.shape[1] for c in y]:
        raise ValueError('Dimensions do not match.')
    out = [core_product(a, b) for a, b in zip(x, y)]
    if all(c.shape[1] == 1 for c in out) and all(c.shape[2] == 1 for c in out):
        v = out[0].reshape(out[0].shape[0], -1)
        for c in out[1:]:
            v = v @ c.reshape(c.shape[0], -1)
        return v[0, 0]
    return out


def matricize(cores):
    """Dense matrix (operator) or vector (tensor) (tensor_train.py:1061-1090)."""
    m = int(np.prod([c.shape[1] for c in cores]))
    n = int(np.prod([c.shape[2] for c in cores]))
    t = cores[0].reshape(cores[0].shape[1], cores[0].shape[2], cores[0].shape[3])
    for c in cores[1:]:
        # t[(M), (N), r] x c[r, m, n, r'] -> [(M m), (N n), r']
        u = np.tensordot(t, c, axes=(2, 0))          # M, N, m, n, r'
        u = u.transpose(0, 2, 1, 3, 4)
        t = u.reshape(u.shape[0] * u.shape[1], u.shape[2] * u.shape[3], u.shape[4])
    return t.reshape(m) if n == 1 else t.reshape(m, n)


def truncated_svd(mat, threshold=0.0, max_rank=np.inf, rel_truncation=True):
    """Thin SVD + truncation (utils.py:111-158).

    gesdd with gesvd retry; keep ``s/s[0] > threshold`` (strict) when the
    threshold is non-zero, then the first ``max_rank`` triplets.
    """
    try:
        u, s, v = sla.svd(mat, full_matrices=False, check_finite=False)
    except sla.LinAlgError:
        u, s, v = sla.svd(mat, full_matrices=False, check_finite=False,
                          lapack_driver='gesvd')
    if threshold != 0:
        keep = (s / s[0] > threshold) if rel_truncation else (s > threshold)
        idx = np.nonzero(keep)[0]
        u, s, v = u[:, idx], s[idx], v[idx, :]
    if max_rank != np.inf:
        k = int(min(len(s), max_rank))
        u, s, v = u[:, :k], s[:k], v[:k, :]
    return u, s, v


def _max_rank_list(order, max_rank):
    """max_rank argument normalisation (tensor_train.py:1146-1159)."""
    if (isinstance(max_rank, (int, np.integer)) and max_rank > 0) or max_rank == np.inf:
        return [1] + [max_rank] * (order - 1) + [1]
    if isinstance(max_rank, list) and len(max_rank) == order + 1 and all(
            (isinstance(r, (int, np.integer)) and r > 0) or r == np.inf for r in max_rank):
        return list(max_rank)
    raise ValueError('Maximum rank(s) must be positive integers.')


def ortho_left(cores, start_index=0, end_index=None, threshold=0.0, max_rank=np.inf):
    """Left-orthonormalisation sweep (tensor_train.py:1092-1205).

    For i = start..end: SVD of core i as (r_i*m_i*n_i, r_{i+1}), truncate,
    keep U as core i, push S*V^T into core i+1.
    """
    cores = copy(cores)
    d = len(cores)
    if end_index is None:
        end_index = d - 2
    if not (isinstance(start_index, int) and isinstance(end_index, int)):
        raise TypeError('Start and end indices must be integers.')
    if not (isinstance(threshold, (int, float)) and threshold >= 0):
        raise ValueError('Threshold must be greater or equal 0.')
    caps = _max_rank_list(d, max_rank)
    for i in range(start_index, end_index + 1):
        r0, m, n, r1 = cores[i].shape
        u, s, v = _svd_trunc(cores[i].reshape(r0 * m * n, r1), threshold, caps[i + 1])
        k = u.shape[1]
        cores[i] = u.reshape(r0, m, n, k)
        cores[i + 1] = np.tensordot(s[:, None] * v, cores[i + 1], axes=(1, 0))
    return cores


def ortho_right(cores, start_index=None, end_index=1, threshold=0.0, max_rank=np.inf):
    """Right-orthonormalisation sweep (tensor_train.py:1207-1310).

    For i = start..end (descending): SVD of core i as (r_i, m_i*n_i*r_{i+1}),
    truncate, keep V^T as core i, push U*S into core i-1.
    """
    cores = copy(cores)
    d = len(cores)
    if start_index is None:
        start_index = d - 1
    if not (isinstance(start_index, (int, np.integer)) and isinstance(end_index, (int, np.integer))):
        raise TypeError('Start and end indices must be integers.')
    if not (isinstance(threshold, (int, float, np.floating, np.integer)) and threshold >= 0):
        raise ValueError('Threshold must be greater or equal 0.')
    caps = _max_rank_list(d, max_rank)
    for i in range(start_index, end_index - 1, -1):
        r0, m, n, r1 = cores[i].shape
        u, s, v = _svd_trunc(cores[i].reshape(r0, m * n * r1), threshold, caps[i])
        k = v.shape[0]
        cores[i] = v.reshape(k, m, n, r1)
        p0, pm, pn, _ = cores[i - 1].shape
        cores[i - 1] = (cores[i - 1].reshape(p0 * pm * pn, -1) @ (u * s[None, :])).reshape(p0, pm, pn, k)
    return cores


def _svd_trunc(mat, threshold, cap):
    """SVD step shared by ortho_left/right (tensor_train.py:1162-1182)."""
    u, s, v = truncated_svd(mat, threshold, np.inf)
    if cap != np.inf:
        k = int(min(len(s), cap))
        u, s, v = u[:, :k], s[:k], v[:k, :]
    return u, s, v


def ortho(cores, threshold=0.0, max_rank=np.inf):
    """ortho_left(thr, inf) then ortho_right(thr, max_rank) (tensor_train.py:1312-1332)."""
    return ortho_right(ortho_left(cores, threshold=threshold, max_rank=np.inf),
                       threshold=threshold, max_rank=max_rank)


def norm(cores, p=2):
    """TT norm (tensor_train.py:1334-1430).

    p=2: vectorise operator cores, right-orthonormalise, Frobenius norm of
    core 0.  p=1: sum over the row modes, matricise, max (non-negative
    entries assumed, as in the reference).
    """
    if p == 2:
        t = [c.reshape(c.shape[0], c.shape[1] * c.shape[2], 1, c.shape[3]) for c in cores]
        t = ortho_right(t)
        return float(np.linalg.norm(t[0].ravel()))
    if p == 1:
        t = copy(cores)
        if all(c.shape[1] == 1 for c in t):
            t = [c.transpose(0, 2, 1, 3) for c in t]
        t = [c.sum(axis=1, keepdims=True) for c in t]
        return float(np.max(matricize(t)))
    raise ValueError('p must be 1 or 2.')


def residual_error(op, x, b):
    """||A x - b|| without forming A x - b (tensor_train.py:2035-2074).

    Per core the block core [A x | b] is formed and SVD-compressed from left
    to right; the norm of the last contracted core is returned.
    """
    d = len(op)
    carry = None
    err = None
    for i in range(d):
        a = op[i]
        xc = x[i]
        R0, m, n, R1 = a.shape
        r0, _, _, r1 = xc.shape
        ax = np.tensordot(a, xc, axes=(2, 1))            # R0 m R1 r0 1 r1
        ax = ax.transpose(0, 3, 1, 4, 2, 5).reshape(R0 * r0, m, R1 * r1)
        bc = b[i].reshape(b[i].shape[0], b[i].shape[1], b[i].shape[3])
        if i == 0:
            core = np.concatenate([ax, -bc], axis=2)
        elif i == d - 1:
            core = np.concatenate([ax, bc], axis=0)
            core = np.tensordot(carry, core, axes=(1, 0))
            err = float(np.linalg.norm(core.ravel()))
            break
        else:
            top = np.concatenate([ax, np.zeros((ax.shape[0], m, bc.shape[2]))], axis=2)
            bot = np.concatenate([np.zeros((bc.shape[0], m, ax.shape[2])), bc], axis=2)
            core = np.tensordot(carry, np.concatenate([top, bot], axis=0), axes=(1, 0))
        u, s, v = sla.svd(core.reshape(core.shape[0] * core.shape[1], core.shape[2]),
                          full_matrices=False, check_finite=False, lapack_driver='gesvd')
        carry = s[:, None] * v
    return err


def eye(dims):
    """tensor_train.py:1870-1893"""
    return [np.eye(n).reshape(1, n, n, 1) for n in dims]


def tt_svd(full, threshold=0.0, max_rank=np.inf):
    """TT decomposition of a full operator/tensor (tensor_train.py:200-261).

    ``full`` has shape (m_1..m_d, n_1..n_d); modes are interleaved
    (m_i, n_i) and split left to right by thin SVDs with the relative
    threshold s/s0 > thr and max_rank.
    """
    d = full.ndim // 2
    if full.ndim % 2:
        raise ValueError('Number of dimensions must be a multiple of 2.')
    m, n = full.shape[:d], full.shape[d:]
    perm = [k for i in range(d) for k in (i, d + i)]
    y = np.transpose(full, perm).copy()
    r = 1
    cores = []
    for i in range(d - 1):
        y = y.reshape(r * m[i] * n[i], -1)
        u, s, v = sla.svd(y, full_matrices=False, check_finite=False)
        if threshold != 0:
            keep = np.nonzero(s / s[0] > threshold)[0]
            u, s, v = u[:, keep], s[keep], v[keep, :]
        if max_rank != np.inf:
            k = int(min(len(s), max_rank))
            u, s, v = u[:, :k], s[:k], v[:k, :]
        cores.append(u.reshape(r, m[i], n[i], u.shape[1]))
        r = u.shape[1]
        y = s[:, None] * v
    cores.append(y.reshape(r, m[-1], n[-1], 1))
    return cores
