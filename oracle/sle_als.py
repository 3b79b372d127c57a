"""Oracle: ALS for A x = b in TT format, dense local solves (TEST ONLY).

Restates scikit_tt/solvers/sle.py:10-95 (driver), 194-305 (interface
stacks), 308-347 (micro matrix), 393-430 (micro rhs) and 475-547 (core
update).  Index conventions:

* op stack ``(x-rank, op-rank, xbar-rank)``: x meets the operator column
  index, xbar the row index (sle.py:217-219, 274-276);
* rhs stack ``(rhs-rank, xbar-rank)`` (sle.py:246-247, 303-305);
* micro matrix rows ``(xbar_L, m, xbar_R)``, cols ``(x_L, n, x_R)``
  (sle.py:343-345).
"""
import numpy as np
import scipy.linalg as sla


def stack_left_op(prev, xc, ac):
    """L_{i+1}[b,Q,b'] = sum L_i[a,P,a'] x[a,n,b] A[P,m,n,Q] conj(x)[a',m,b'] (sle.py:214-219)."""
    x3 = xc[:, :, 0, :]
    t = np.tensordot(prev, x3, axes=(0, 0))               # P a' n b
    t = np.tensordot(t, ac, axes=([0, 2], [0, 2]))        # a' b m Q
    return np.tensordot(t, np.conj(x3), axes=([0, 2], [0, 1]))   # b Q b'


def stack_right_op(nxt, xc, ac):
    """R_{i-1}[a,P,a'] = sum x[a,n,b] A[P,m,n,Q] R_i[b,Q,b'] conj(x)[a',m,b'] (sle.py:271-276)."""
    x3 = xc[:, :, 0, :]
    t = np.tensordot(np.conj(x3), nxt, axes=(2, 2))       # a' m b Q
    t = np.tensordot(ac, t, axes=([1, 3], [1, 3]))        # P n a' b
    return np.tensordot(x3, t, axes=([1, 2], [1, 3]))     # a P a'


def stack_left_rhs(prev, xc, bc):
    """Lb_{i+1}[q,b'] = sum Lb_i[p,a'] B[p,m,q] conj(x)[a',m,b'] (sle.py:242-247)."""
    t = np.tensordot(prev, bc[:, :, 0, :], axes=(0, 0))   # a' m q
    return np.tensordot(t, np.conj(xc[:, :, 0, :]), axes=([0, 1], [0, 1]))


def stack_right_rhs(nxt, xc, bc):
    """Rb_{i-1}[p,a'] = sum B[p,m,q] Rb_i[q,b'] conj(x)[a',m,b'] (sle.py:300-305)."""
    t = np.tensordot(np.conj(xc[:, :, 0, :]), nxt, axes=(2, 1))   # a' m q
    return np.tensordot(bc[:, :, 0, :], t, axes=([1, 2], [1, 2]))  # p a'


def micro_matrix(L, ac, R):
    """Dense local operator (sle.py:338-347): rows (a',m,b'), cols (a,n,b)."""
    t = np.tensordot(L, ac, axes=(1, 0))                  # a a' m n Q
    t = np.tensordot(t, R, axes=(4, 1))                   # a a' m n b b'
    a, a2, m, n, b, b2 = t.shape
    return t.transpose(1, 2, 5, 0, 3, 4).reshape(a2 * m * b2, a * n * b)


def micro_rhs(Lb, bc, Rb):
    """Local right-hand side (sle.py:424-430), rows (a',m,b')."""
    t = np.tensordot(Lb, bc[:, :, 0, :], axes=(0, 0))     # a' m q
    t = np.tensordot(t, Rb, axes=(2, 0))                  # a' m b'
    return t.reshape(-1, 1)


def micro_matvec(L, ac, R, v):
    """y = M v without forming M (same contraction as micro_matrix)."""
    a, P, a2 = L.shape
    b, Q, b2 = R.shape
    n = ac.shape[2]
    v3 = v.reshape(a, n, b)
    t = np.tensordot(L, v3, axes=(0, 0))                  # P a' n b
    t = np.tensordot(t, ac, axes=([0, 2], [0, 2]))        # a' b m Q
    t = np.tensordot(t, R, axes=([1, 3], [0, 1]))         # a' m b'
    return t.reshape(-1)


def update_core(i, M, f, cores, ranks, solver, direction):
    """Local solve + orthonormalisation (sle.py:475-547).

    'solve' -> LAPACK gesv, 'lu' -> getrf/getrs; an unknown solver leaves the
    previous core in place, as the reference does (sle.py:505-509).  Forward:
    economic QR of (r_i*n_i, r_{i+1}), keep Q.  Backward, i>0: economic RQ of
    (r_i, n_i*r_{i+1}), keep Q.  Backward, i=0: keep the raw solution.
    """
    n_i = cores[i].shape[1]
    vec = cores[i]
    if solver == 'solve':
        vec = np.linalg.solve(M, f)
    if solver == 'lu':
        lu = sla.lu_factor(M, check_finite=False)
        vec = sla.lu_solve(lu, f, check_finite=False)
    if direction == 'forward':
        q, _ = sla.qr(np.reshape(vec, (ranks[i] * n_i, ranks[i + 1])), mode='economic', check_finite=False)
        ranks[i + 1] = q.shape[1]
        cores[i] = q.reshape(ranks[i], n_i, 1, ranks[i + 1])
    else:
        if i > 0:
            _, q = sla.rq(np.reshape(vec, (ranks[i], n_i * ranks[i + 1])), mode='economic', check_finite=False)
            ranks[i] = q.shape[0]
            cores[i] = q.reshape(ranks[i], n_i, 1, ranks[i + 1])
        else:
            cores[i] = np.reshape(vec, (ranks[i], n_i, 1, ranks[i + 1]))


def als(operator, initial_guess, right_hand_side, repeats=1, solver='solve'):
    """ALS driver (sle.py:10-95).

    The right stacks are built from the (not re-orthonormalised) copy of the
    initial guess; each repeat is a forward half sweep over i=0..d-2 and a
    backward half sweep over i=d-1..0.
    """
    d = len(operator)
    cores = [np.array(c, copy=True) for c in initial_guess]
    ranks = [c.shape[0] for c in cores] + [cores[-1].shape[3]]
    one3 = np.ones((1, 1, 1))
    one2 = np.ones((1, 1))
    Lop, Lb = [None] * d, [None] * d
    Rop, Rb = [None] * d, [None] * d

    def right(i):
        if i == d - 1:
            Rop[i], Rb[i] = one3, one2
        else:
            Rop[i] = stack_right_op(Rop[i + 1], cores[i + 1], operator[i + 1])
            Rb[i] = stack_right_rhs(Rb[i + 1], cores[i + 1], right_hand_side[i + 1])

    def left(i):
        if i == 0:
            Lop[i], Lb[i] = one3, one2
        else:
            Lop[i] = stack_left_op(Lop[i - 1], cores[i - 1], operator[i - 1])
            Lb[i] = stack_left_rhs(Lb[i - 1], cores[i - 1], right_hand_side[i - 1])

    def step(i, direction):
        M = micro_matrix(Lop[i], operator[i], Rop[i])
        f = micro_rhs(Lb[i], right_hand_side[i], Rb[i])
        update_core(i, M, f, cores, ranks, solver, direction)

    for i in range(d - 1, -1, -1):
        right(i)
    for _ in range(repeats):
        for i in range(d):
            left(i)
            if i < d - 1:
                step(i, 'forward')
        for i in range(d - 1, -1, -1):
            right(i)
            step(i, 'backward')
    return cores
