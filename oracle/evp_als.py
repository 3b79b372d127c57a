"""Oracle: ALS for (generalised) eigenproblems in TT format (TEST ONLY).

Restates scikit_tt/solvers/evp.py:17-179 (driver), 253-334 (stacks),
337-383 (micro matrices) and 386-495 (core update) for ``previous=[]``.
For real data the evp stacks are numerically identical to the sle ones
(evp.py:281-283 vs sle.py:217-219), so the sle oracle's stack functions are
reused.
"""
import numpy as np
import scipy.linalg as sla

from .sle_als import stack_left_op, stack_right_op, micro_matrix


def local_eigen(M, B, k, solver, sigma, real):
    """Local eigenproblem and selection rule (evp.py:411-443).

    'eig': all eigenpairs (dgeev/dggev), the k nearest ``sigma`` in ascending
    distance.  'eigh': the k largest (dsyevr-class), largest first.
    """
    vals = vecs = None
    if solver == 'eig':
        vals, vecs = sla.eig(M, b=B, check_finite=False)
        order = np.argsort(np.abs(vals - sigma))
        vals, vecs = vals[order[:k]], vecs[:, order[:k]]
    elif solver == 'eigh':
        N = M.shape[0]
        vals, vecs = sla.eigh(M, b=B, check_finite=False, subset_by_index=(N - k, N - 1))
        vals, vecs = vals[::-1], vecs[:, ::-1]
    else:
        raise NotImplementedError(solver)
    if real:
        vals = np.real(vals)
    return vals, vecs


def update_core(i, vecs, cores, ranks, k, direction):
    """Orthonormalisation after the local eigensolve (evp.py:445-493).

    Forward: full SVD (gesvd) of (r_i*n_i, r_{i+1}*k), keep the first
    min(r_{i+1}, r_i*n_i) left singular vectors.  Backward, i>0: full SVD of
    vecs^T as (k*r_i, n_i*r_{i+1}), keep the first min(r_i, n_i*r_{i+1}) right
    singular vectors.  Backward, i=0: core 0 becomes (r0, n0, 1, r1, k).
    """
    n_i = cores[i].shape[1]
    if direction == 'forward':
        u, _, _ = sla.svd(vecs.reshape(ranks[i] * n_i, ranks[i + 1] * k), lapack_driver='gesvd',
                          check_finite=False)
        r = min(ranks[i + 1], u.shape[1])
        ranks[i + 1] = r
        cores[i] = u[:, :r].reshape(ranks[i], n_i, 1, r)
    elif i > 0:
        _, _, v = sla.svd(vecs.T.reshape(k * ranks[i], n_i * ranks[i + 1]), lapack_driver='gesvd',
                          check_finite=False)
        r = min(ranks[i], v.shape[0])
        ranks[i] = r
        cores[i] = v[:r, :].reshape(r, n_i, 1, ranks[i + 1])
    else:
        cores[i] = vecs.reshape(ranks[i], n_i, 1, ranks[i + 1], k)


def als(operator, initial_guess, operator_gevp=None, number_ev=1, repeats=1, conv_eps=1e-10,
        solver='eig', sigma=1, real=True):
    """evp.als driver (evp.py:17-179) for ``previous=[]``.

    Returns ``(eigenvalue, cores, iterations)`` for number_ev == 1 with the
    best-so-far pair by |lambda - sigma| (evp.py:151-155), else the list of
    eigenvalues and the list of k eigentensors.  Convergence: max |dlambda|
    over the last <= 3 sweeps (history seeded with +inf) below conv_eps
    (evp.py:158-165).
    """
    d = len(operator)
    k = number_ev
    cores = [np.array(c, copy=True) for c in initial_guess]
    ranks = [c.shape[0] for c in cores] + [cores[-1].shape[3]]
    one3 = np.ones((1, 1, 1))
    L, R = [None] * d, [None] * d
    Lg, Rg = [None] * d, [None] * d

    def right(i):
        if i == d - 1:
            R[i] = Rg[i] = one3
        else:
            R[i] = stack_right_op(R[i + 1], cores[i + 1], operator[i + 1])
            if operator_gevp is not None:
                Rg[i] = stack_right_op(Rg[i + 1], cores[i + 1], operator_gevp[i + 1])

    def left(i):
        if i == 0:
            L[i] = Lg[i] = one3
        else:
            L[i] = stack_left_op(L[i - 1], cores[i - 1], operator[i - 1])
            if operator_gevp is not None:
                Lg[i] = stack_left_op(Lg[i - 1], cores[i - 1], operator_gevp[i - 1])

    def step(i, direction):
        M = micro_matrix(L[i], operator[i], R[i])
        B = None if operator_gevp is None else micro_matrix(Lg[i], operator_gevp[i], Rg[i])
        vals, vecs = local_eigen(M, B, k, solver, sigma, real)
        update_core(i, vecs, cores, ranks, k, direction)
        return vals

    for i in range(d - 1, -1, -1):
        right(i)
    history = np.full((1, k), np.inf)
    best_val, best_cores = np.inf, None
    it = 1
    converged = False
    vals = None
    while it <= repeats and not converged:
        for i in range(d):
            left(i)
            if i < d - 1:
                vals = step(i, 'forward')
        for i in range(d - 1, -1, -1):
            right(i)
            vals = step(i, 'backward')
        it += 1
        if k == 1 and np.abs(vals[0] - sigma) < np.abs(best_val - sigma):
            best_val = vals[0].copy()
            best_cores = [cores[0][:, :, :, :, 0]] + list(cores[1:])
        recent = history[-min(3, history.shape[0]):, :]
        if np.amax(np.abs(recent - vals)) < conv_eps:
            converged = True
        history = np.vstack((history, vals))
    if k == 1:
        return best_val, best_cores, it - 1
    return vals, [[cores[0][:, :, :, :, j]] + list(cores[1:]) for j in range(k)], it - 1
