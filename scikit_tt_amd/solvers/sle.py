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
from .. import _lib
from ..tensor_train import TT, _stream
from . import _engine

__all__ = ['als', 'als_batch', 'als_gang']

_SOLVERS = {'solve': 0, 'lu': 1}


def als_batch(operators, initial_guesses, right_hand_sides, repeats=1, solver='solve', tol=1e-12, maxit=4000,
              return_info=False):
    """Solve independent systems A_s x_s = b_s in ONE engine launch.

    Returns a sequence of solutions (and a sequence of SolveInfo if
    ``return_info``), each made on access.  A singular local LU raises numpy.linalg.LinAlgError
    (reported for the first failing system).
    """
    if solver not in _SOLVERS:
        raise NotImplementedError("solver must be 'solve' or 'lu', got %r" % (solver,))
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
    rc = _lib.load().ttals_sle_als_batch(ctx, _stream(dev), count, A, X, B, O, int(repeats), _SOLVERS[solver],
                                         float(tol), int(maxit), res)
    _lib.check(rc, ctx, 'sle.als')
    infos = _engine.Infos(res)
    if not return_info:
        _engine.raise_first(infos, 'sle.als', count)
    sols = _engine.Solutions(cores, opacks, initial_guesses)
    return (sols, infos) if return_info else sols


_PRECISIONS = {'fp64': 0, 'mixed': 1}


def als_gang(operators, initial_guesses, right_hand_sides, repeats=1, solver='solve', precision='fp64', gang=0,
             tol=1e-12, maxit=4000, return_info=False):
    """Large systems on the multi-workgroup engine: each system is solved by a
    gang of ``gang`` workgroups (0: one per row of its widest bond, up to 32)
    that split the local matvec, the interface stacks and the Krylov vector
    operations; all gangs run at once in one cooperative launch.

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


def als(operator, initial_guess, right_hand_side, repeats=1, solver='solve'):
    """Alternating linear scheme (scikit_tt/solvers/sle.py:10-95) on the device.

    Deviation from the reference: an unknown ``solver`` raises
    NotImplementedError (the reference silently keeps the previous core).
    """
    _engine.check_tts([operator, initial_guess, right_hand_side])
    return als_batch([operator], [initial_guess], [right_hand_side], repeats=repeats, solver=solver)[0]
