"""Alternating linear scheme for A x = b in TT format on the HIP device.

Drop-in for scikit_tt.solvers.sle.als (scikit_tt/solvers/sle.py:10-95): same
signature, defaults and return convention (a new TT with cores
``(r_i, n_i, 1, r_{i+1})``, cores 1..d-1 right-orthonormal, the norm in core
0).  One call is one launch of the fused engine in libttals.so; ``als_batch``
solves many independent systems of uneven order and rank in one launch.

Local solves: LU with partial pivoting (as dgesv, sle.py:506) for local size
<= 96, matrix-free CG (operator symmetric slice by slice; falls back to
GMRES if p'Mp <= 0) or GMRES(40) above, to relative residual ``tol``.
The initial guess is right-orthonormalised on the device before the right
stacks are built; ALS is gauge invariant, so this changes nothing in exact
arithmetic and keeps the stacks well scaled.
"""
import ctypes

import numpy as np

from .. import _lib
from ..tensor_train import TT, _stream
from . import _engine

__all__ = ['als', 'mals', 'als_batch', 'als_gang']

_SOLVERS = {'solve': 0, 'lu': 1}


_PRECISIONS = {'fp64': 0, 'mixed': 1}
_SCHEDULES = {'one': 0, 'auto': 1}


def als_batch(operators, initial_guesses, right_hand_sides, repeats=1, solver='solve', tol=1e-12, maxit=4000,
              return_info=False, schedule='one', gang_precision='fp64'):
    """Solve independent systems A_s x_s = b_s in ONE engine launch.

    ``schedule='one'``: one workgroup per system.  ``schedule='auto'``: the
    systems that the device's launch-time model (``scikit_tt_amd.tune``)
    says would set the launch time are solved by gangs of workgroups (as
    ``als_gang``, in ``gang_precision``), the rest by one workgroup each, all
    in the same launch; ``infos.gangs`` then holds each system's gang size
    (0: one workgroup).

    Returns a sequence of solutions (and a sequence of SolveInfo if
    ``return_info``), each made on access.  A singular local LU raises numpy.linalg.LinAlgError
    (reported for the first failing system).
    """
    if solver not in _SOLVERS:
        raise NotImplementedError("solver must be 'solve' or 'lu', got %r" % (solver,))
    if schedule not in _SCHEDULES:
        raise ValueError("schedule must be 'one' or 'auto', got %r" % (schedule,))
    if gang_precision not in _PRECISIONS:
        raise ValueError("gang_precision must be 'fp64' or 'mixed', got %r" % (gang_precision,))
    count = len(operators)
    if count == 0 or len(initial_guesses) != count or len(right_hand_sides) != count:
        raise ValueError('operators, initial guesses and right-hand sides must be non-empty lists of equal length')
    _engine.check_tts(operators, initial_guesses, right_hand_sides)
    dev = operators[0].device
    dA, A = _engine.descs(operators)
    dX, X = _engine.descs(initial_guesses)
    dB, B = _engine.descs(right_hand_sides)
    cores, opacks, O = _engine.outputs(initial_guesses)
    res = (_lib.Result * count)()
    ctx = _lib.context(dev.index)
    lib = _lib.load()
    with _lib.schedule(ctx, _SCHEDULES[schedule], _PRECISIONS[gang_precision]):
        rc = lib.ttals_sle_als_batch(ctx, _stream(dev), count, A, X, B, O, int(repeats), _SOLVERS[solver],
                                     float(tol), int(maxit), res)
        _lib.check(rc, ctx, 'sle.als')
        gangs = np.zeros(count, dtype=np.int32)
        _lib.check(lib.ttals_last_plan(ctx, gangs.ctypes.data, count), ctx, 'last_plan')
    infos = _engine.Infos(res)
    infos.gangs = gangs
    if not return_info:
        _engine.raise_first(infos, 'sle.als', count)
    sols = _engine.Solutions(cores, opacks, initial_guesses)
    return (sols, infos) if return_info else sols


def als_gang(operators, initial_guesses, right_hand_sides, repeats=1, solver='solve', precision='fp64', gang=0,
             tol=1e-12, maxit=4000, return_info=False):
    """Large systems on the multi-workgroup engine: each system is solved by a
    gang of ``gang`` workgroups (0: one per row of its widest bond, up to 32)
    that split the local matvec, the interface stacks and the Krylov vector
    operations; all gangs run at once in one launch.

    ``precision='mixed'``: Krylov iterations in FP32 on an FP32 copy of the
    local operator, refined with the FP64 local residual until it meets
    ``tol`` (cores, stacks, QR and the residual FP64).  Results agree with
    ``precision='fp64'`` to the local tolerance.  Same return convention as
    ``als_batch``.
    """
    if solver not in _SOLVERS:
        raise NotImplementedError("solver must be 'solve' or 'lu', got %r" % (solver,))
    if precision not in _PRECISIONS:
        raise ValueError("precision must be 'fp64' or 'mixed', got %r" % (precision,))
    count = len(operators)
    if count == 0 or len(initial_guesses) != count or len(right_hand_sides) != count:
        raise ValueError('operators, initial guesses and right-hand sides must be non-empty lists of equal length')
    _engine.check_tts(operators, initial_guesses, right_hand_sides)
    dev = operators[0].device
    dA, A = _engine.descs(operators)
    dX, X = _engine.descs(initial_guesses)
    dB, B = _engine.descs(right_hand_sides)
    cores, opacks, O = _engine.outputs(initial_guesses)
    res = (_lib.Result * count)()
    ctx = _lib.context(dev.index)
    rc = _lib.load().ttals_sle_als_gang(ctx, _stream(dev), count, A, X, B, O, int(repeats), _SOLVERS[solver],
                                        float(tol), int(maxit), _PRECISIONS[precision], int(gang), res)
    _lib.check(rc, ctx, 'sle.als (gang)')
    infos = _engine.Infos(res)
    if not return_info:
        _engine.raise_first(infos, 'sle.als', count)
    sols = _engine.Solutions(cores, opacks, initial_guesses)
    return (sols, infos) if return_info else sols


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def mals(operator, initial_guess, right_hand_side, repeats=1, solver='solve', threshold=1e-12, max_rank=np.inf):
    """Modified ALS, two-site and rank-adaptive (scikit_tt/solvers/sle.py:98-191).

    Same sweep as the reference: left stacks forward, right stacks
    backward; the two-site problem at (i, i+1) (stacks L_i, R_{i+1} and the
    supercore of operator cores i, i+1, sle.py:350-390, 433-472) is solved on
    the device, the two-site solution split by a truncated SVD
    (s/s0 > threshold, then max_rank, sle.py:589-645): forward keeps U in
    core i, backward V^T in core i+1 and U S in core 0 at the end.  Ranks
    grow and shrink with the truncation, so the sweep runs from the host,
    piece by piece (stack updates, supercores, local rhs, local solve,
    truncated SVD are device calls); local systems of size <= 96 use LU
    with partial pivoting like the reference, larger ones CG/GMRES to
    relative residual 1e-12 (as in ``als``).
    """
    import torch
    if solver not in _SOLVERS:
        raise NotImplementedError("solver must be 'solve' or 'lu', got %r" % (solver,))
    _engine.check_tts([operator, initial_guess, right_hand_side])
    from ..utils import truncated_svd
    lib = _lib.load()
    dev = operator.device
    ctx = _lib.context(dev.index)
    st = _stream(dev)
    d = operator.order
    n = list(operator.row_dims)
    ra, rb = list(operator.ranks), list(right_hand_side.ranks)
    A = [c.contiguous() for c in operator.device_cores]
    B = [c.contiguous() for c in right_hand_side.device_cores]
    ranks = list(initial_guess.ranks)
    cores = [c.reshape(c.shape[0], c.shape[1], c.shape[3]).clone() for c in initial_guess.device_cores]
    sym = all(bool(torch.equal(c, c.transpose(1, 2))) for c in A)
    f64 = dict(dtype=torch.float64, device=dev)
    Lop, Lb, Rop, Rb = [None] * d, [None] * d, [None] * d, [None] * d

    def chk(rc, what):
        _lib.check(rc, ctx, what)

    def left(i):
        if i == 0:
            Lop[0], Lb[0] = torch.ones((1, 1, 1), **f64), torch.ones((1, 1), **f64)
            return
        j = i - 1
        a, b = ranks[j], ranks[i]
        Lop[i] = torch.empty((b, ra[i], b), **f64)
        chk(lib.ttals_stack_left_op(ctx, st, a, ra[j], n[j], ra[i], b, _ptr(Lop[j]), _ptr(A[j]), _ptr(cores[j]),
                                    _ptr(Lop[i])), 'stack_left_op')
        Lb[i] = torch.empty((rb[i], b), **f64)
        chk(lib.ttals_stack_left_rhs(ctx, st, a, rb[j], n[j], rb[i], b, _ptr(Lb[j]), _ptr(B[j]), _ptr(cores[j]),
                                     _ptr(Lb[i])), 'stack_left_rhs')

    def right(i):
        if i == d - 1:
            Rop[i], Rb[i] = torch.ones((1, 1, 1), **f64), torch.ones((1, 1), **f64)
            return
        j = i + 1
        a, b = ranks[j], ranks[j + 1]
        Rop[i] = torch.empty((a, ra[j], a), **f64)
        chk(lib.ttals_stack_right_op(ctx, st, a, ra[j], n[j], ra[j + 1], b, _ptr(Rop[j]), _ptr(A[j]),
                                     _ptr(cores[j]), _ptr(Rop[i])), 'stack_right_op')
        Rb[i] = torch.empty((rb[j], a), **f64)
        chk(lib.ttals_stack_right_rhs(ctx, st, a, rb[j], n[j], rb[j + 1], b, _ptr(Rb[j]), _ptr(B[j]),
                                      _ptr(cores[j]), _ptr(Rb[i])), 'stack_right_rhs')

    def merge(X, Y, r1, m1, n1, k, m2, n2, r2):
        out = torch.empty((r1, m1 * m2, n1 * n2, r2), **f64)
        rc = lib.ttals_core_merge(st, r1, m1, n1, k, m2, n2, r2, _ptr(X), _ptr(Y), _ptr(out))
        if rc != 0:
            raise RuntimeError('core_merge failed (%d)' % rc)
        return out

    def step(i, direction):
        a, b = ranks[i], ranks[i + 2]
        n1, n2 = n[i], n[i + 1]
        nn = n1 * n2
        N = a * nn * b
        Asup = merge(A[i], A[i + 1], ra[i], n1, n1, ra[i + 1], n2, n2, ra[i + 2])
        Bsup = merge(B[i], B[i + 1], rb[i], n1, 1, rb[i + 1], n2, 1, rb[i + 2])
        f = torch.empty(N, **f64)
        chk(lib.ttals_micro_rhs(ctx, st, a, rb[i], nn, rb[i + 2], b, _ptr(Lb[i]), _ptr(Bsup), _ptr(Rb[i + 1]),
                                _ptr(f)), 'micro_rhs')
        # warm start (matrix-free local solves): the current two-site contraction
        v = torch.tensordot(cores[i], cores[i + 1], dims=([2], [0])).reshape(N).contiguous()
        its = ctypes.c_int()
        rc = lib.ttals_local_solve(ctx, st, a, ra[i], nn, ra[i + 2], b, _ptr(Lop[i]), _ptr(Asup), _ptr(Rop[i + 1]),
                                   _ptr(f), _ptr(v), 0, 1 if sym else 0, 1e-12, 4000, ctypes.byref(its))
        if rc == _lib.S_SINGULAR:
            raise np.linalg.LinAlgError('Singular matrix (sle.mals)')
        chk(rc if rc < 0 else 0, 'local_solve')
        u, s, vt = truncated_svd(v.reshape(a * n1, n2 * b), threshold=threshold, max_rank=max_rank)
        k = s.shape[0]
        ranks[i + 1] = k
        if direction == 'forward':
            cores[i] = u.reshape(a, n1, k).contiguous()
            cores[i + 1] = (s[:, None] * vt).reshape(k, n2, b).contiguous()
        else:
            cores[i + 1] = vt.reshape(k, n2, b).contiguous()
            cores[i] = (u * s[None, :]).reshape(a, n1, k).contiguous()

    for i in range(d - 1, 0, -1):
        right(i)
    for _ in range(repeats):
        for i in range(d - 1):
            left(i)
            if i < d - 2:
                step(i, 'forward')
        for i in range(d - 2, -1, -1):
            right(i + 1)
            step(i, 'backward')
    return TT._from_device([c.reshape(c.shape[0], c.shape[1], 1, c.shape[2]) for c in cores])


def als(operator, initial_guess, right_hand_side, repeats=1, solver='solve'):
    """Alternating linear scheme (scikit_tt/solvers/sle.py:10-95) on the device.

    Deviation from the reference: an unknown ``solver`` raises
    NotImplementedError (the reference silently keeps the previous core).
    """
    _engine.check_tts([operator, initial_guess, right_hand_side])
    return als_batch([operator], [initial_guess], [right_hand_side], repeats=repeats, solver=solver)[0]
