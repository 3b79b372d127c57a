"""Pins of the CPU oracle (no GPU).

Each oracle function is checked against something that does not restate
it: golden vectors written by the reference itself (tests/golden), dense
brute force on tiny inputs, or closed forms.  `test_planted_*` show that a
plausible mistake (transposed stack, wrong QR side, dropped term) fails
those checks.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import tt_ops, sle_als, evp_als
from scikit_tt_amd import workloads as W


def rand_tt(rng, dims, ranks, cols=None):
    return W.random_tt(rng, dims, ranks, col_dims=cols)


def dense_interface_left(cores, i):
    """X_{<i} as a (prod n_<i, r_i) matrix."""
    m = np.ones((1, 1))
    for c in cores[:i]:
        r0, n, _, r1 = c.shape
        m = (m @ c.reshape(r0, n * r1)).reshape(-1, r1)
    return m


def dense_interface_right(cores, i):
    """X_{>i} as a (r_{i+1}, prod n_>i) matrix."""
    m = np.ones((1, 1))
    for c in reversed(cores[i + 1:]):
        r0, n, _, r1 = c.shape
        m = (c.reshape(r0 * n, r1) @ m).reshape(r0, -1)
    return m


def dense_op_partial_left(op, i):
    """A_{<i} as a tensor (M, N, R_i) of the first i operator cores."""
    t = np.ones((1, 1, 1))
    for c in op[:i]:
        u = np.einsum('abr,rmns->ambns', t, c)
        t = u.reshape(u.shape[0] * u.shape[1], u.shape[2] * u.shape[3], u.shape[4])
    return t


def dense_op_partial_right(op, i):
    t = np.ones((1, 1, 1))
    for c in reversed(op[i + 1:]):
        u = np.einsum('rmns,sab->rmanb', c, t)
        t = u.reshape(u.shape[0], u.shape[1] * u.shape[2], u.shape[3] * u.shape[4])
    return t


# --------------------------------------------------------------------------
# TT algebra vs dense


def test_add_sub_scale_dot_matricize():
    rng = np.random.default_rng(0)
    A = rand_tt(rng, [2, 3, 2], [1, 2, 3, 1], cols=[3, 2, 2])
    x = rand_tt(rng, [3, 2, 2], [1, 2, 2, 1])
    y = rand_tt(rng, [3, 2, 2], [1, 3, 1, 1])
    Xd, Yd, Ad = tt_ops.matricize(x), tt_ops.matricize(y), tt_ops.matricize(A)
    assert Ad.shape == (12, 12)
    np.testing.assert_allclose(tt_ops.matricize(tt_ops.add(x, y)), Xd + Yd, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(tt_ops.matricize(tt_ops.sub(x, y)), Xd - Yd, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(tt_ops.matricize(tt_ops.scale(x, -2.5)), -2.5 * Xd, rtol=1e-13)
    np.testing.assert_allclose(tt_ops.matricize(tt_ops.dot(A, x)), Ad @ Xd, rtol=1e-12, atol=1e-12)
    # the product-rank index is i_self * s + i_other (tensor_train.py:471-479)
    p = tt_ops.core_product(A[1], x[1])
    assert np.allclose(p[1 * 2 + 0, :, :, 2 * 2 + 1], A[1][1, :, :, 2] @ x[1][0, :, :, 1])
    # scalar result when all dims are 1
    u = rand_tt(rng, [1, 1], [1, 2, 1])
    v = rand_tt(rng, [1, 1], [1, 3, 1])
    assert np.isclose(tt_ops.dot(u, v), tt_ops.matricize(u) * tt_ops.matricize(v))
    with pytest.raises(ValueError):
        tt_ops.dot(x, x)
    with pytest.raises(ValueError):
        tt_ops.add(x, rand_tt(rng, [3, 2, 3], 2))
    with pytest.raises(TypeError):
        tt_ops.scale(x, 'a')


def test_validate_errors():
    with pytest.raises(ValueError):
        tt_ops.validate([np.zeros((1, 2, 1))])
    with pytest.raises(ValueError):
        tt_ops.validate([np.zeros((1, 2, 1, 2)), np.zeros((3, 2, 1, 1))])
    with pytest.raises(TypeError):
        tt_ops.validate(np.zeros((1, 2, 1, 1)))


def test_norm_and_ortho_vs_dense():
    rng = np.random.default_rng(1)
    x = rand_tt(rng, [3, 4, 2, 3], [1, 3, 5, 2, 1])
    A = rand_tt(rng, [2, 3, 2], 2, cols=[3, 2, 2])
    Xd = tt_ops.matricize(x)
    assert np.isclose(tt_ops.norm(x), np.linalg.norm(Xd), rtol=1e-13)
    assert np.isclose(tt_ops.norm(A), np.linalg.norm(tt_ops.matricize(A), 'fro'), rtol=1e-13)
    xr = tt_ops.ortho_right(x)
    np.testing.assert_allclose(tt_ops.matricize(xr), Xd, rtol=1e-12, atol=1e-12)
    for c in xr[1:]:
        m = c.reshape(c.shape[0], -1)
        np.testing.assert_allclose(m @ m.T, np.eye(m.shape[0]), atol=1e-12)
    # ranks: min(r_i, n_i r_{i+1}) from the right
    assert tt_ops.ranks_of(xr) == [1, 3, 4, 2, 1]
    xl = tt_ops.ortho_left(x)
    np.testing.assert_allclose(tt_ops.matricize(xl), Xd, rtol=1e-12, atol=1e-12)
    for c in xl[:-1]:
        m = c.reshape(-1, c.shape[3])
        np.testing.assert_allclose(m.T @ m, np.eye(m.shape[1]), atol=1e-12)
    assert tt_ops.ranks_of(xl) == [1, 3, 5, 2, 1]
    # p=1 (non-negative entries)
    z = [np.abs(c) for c in x]
    assert np.isclose(tt_ops.norm(z, p=1), np.abs(tt_ops.matricize(z)).sum())


def test_truncation_ranks_exact():
    """s/s0 > thr (strict) then max_rank (utils.py:145-156), on spectra with clear gaps."""
    rng = np.random.default_rng(2)
    U, _ = np.linalg.qr(rng.standard_normal((20, 6)))
    V, _ = np.linalg.qr(rng.standard_normal((9, 6)))
    s = np.array([1.0, 0.5, 1e-3, 1e-6, 1e-9, 1e-12])
    M = (U * s) @ V.T
    # threshold 0 keeps the whole thin SVD (min(20, 9) = 9 triplets)
    for thr, expect in [(0, 9), (1e-2, 2), (1e-4, 3), (1e-7, 4), (1e-10, 5)]:
        u, ss, v = tt_ops.truncated_svd(M, thr)
        assert len(ss) == expect, (thr, len(ss))
        k = min(expect, 5)
        np.testing.assert_allclose(ss[:k], s[:k], rtol=1e-3, atol=1e-15)
    u, ss, v = tt_ops.truncated_svd(M, 1e-10, max_rank=3)
    assert len(ss) == 3
    u, ss, v = tt_ops.truncated_svd(M, 0.6, rel_truncation=False)
    assert len(ss) == 1
    # TT ortho with threshold: a TT whose middle unfolding has a clear gap
    a = rng.standard_normal((1, 4, 1, 6))
    b = rng.standard_normal((6, 4, 1, 1))
    w = np.diag([1, 1e-2, 1e-4, 1e-13, 1e-14, 1e-15])
    x = [a, np.einsum('ij,jmnk->imnk', w, b)]
    assert tt_ops.ranks_of(tt_ops.ortho(x, threshold=1e-8)) == [1, 3, 1]
    assert tt_ops.ranks_of(tt_ops.ortho(x, threshold=1e-8, max_rank=2)) == [1, 2, 1]


def test_residual_error_matches_explicit():
    rng = np.random.default_rng(3)
    A = rand_tt(rng, [2, 3, 2, 2], 2, cols=[2, 3, 2, 2])
    x = rand_tt(rng, [2, 3, 2, 2], 3)
    b = rand_tt(rng, [2, 3, 2, 2], 2)
    dense = np.linalg.norm(tt_ops.matricize(A) @ tt_ops.matricize(x) - tt_ops.matricize(b))
    assert np.isclose(tt_ops.residual_error(A, x, b), dense, rtol=1e-12)
    assert np.isclose(tt_ops.norm(tt_ops.sub(tt_ops.dot(A, x), b)), dense, rtol=1e-12)


def test_tt_svd_reconstructs():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((2, 3, 2, 3, 2, 2))  # operator (2,3,2) x (3,2,2)
    c = tt_ops.tt_svd(X)
    M = X.reshape(12, 12)
    np.testing.assert_allclose(tt_ops.matricize(c), M, rtol=1e-12, atol=1e-12)
    ct = tt_ops.tt_svd(X, threshold=1e-10)
    np.testing.assert_allclose(tt_ops.matricize(ct), M, rtol=1e-9, atol=1e-9)


# --------------------------------------------------------------------------
# stacks and local problems vs dense brute force


@pytest.mark.parametrize('i', [0, 1, 2, 3])
def test_stacks_and_micro_matrix_vs_dense(i):
    rng = np.random.default_rng(10 + i)
    dims = [2, 3, 2, 2]
    A = rand_tt(rng, dims, [1, 2, 3, 2, 1], cols=dims)
    x = rand_tt(rng, dims, [1, 2, 3, 2, 1])
    b = rand_tt(rng, dims, [1, 2, 2, 3, 1])
    d = len(dims)
    # stacks by the oracle recursions
    L = [np.ones((1, 1, 1))]
    Lb = [np.ones((1, 1))]
    for j in range(d - 1):
        L.append(sle_als.stack_left_op(L[-1], x[j], A[j]))
        Lb.append(sle_als.stack_left_rhs(Lb[-1], x[j], b[j]))
    R = [None] * d
    Rb = [None] * d
    R[d - 1], Rb[d - 1] = np.ones((1, 1, 1)), np.ones((1, 1))
    for j in range(d - 2, -1, -1):
        R[j] = sle_als.stack_right_op(R[j + 1], x[j + 1], A[j + 1])
        Rb[j] = sle_als.stack_right_rhs(Rb[j + 1], x[j + 1], b[j + 1])
    # dense: L_i[a, P, a'] = sum_{m,n} X_<i[n, a] A_<i[m, n, P] X_<i[m, a']
    # (x meets the operator column index n, xbar the row index m)
    Xl = dense_interface_left(x, i)
    Al = dense_op_partial_left(A, i)
    Ld = np.einsum('na,mnp,mc->apc', Xl, Al, Xl)
    np.testing.assert_allclose(L[i], Ld, rtol=1e-11, atol=1e-11)
    Xr = dense_interface_right(x, i)
    Ar = dense_op_partial_right(A, i)
    Rd = np.einsum('bn,pmn,cm->bpc', Xr, Ar, Xr)
    np.testing.assert_allclose(R[i], Rd, rtol=1e-11, atol=1e-11)
    Bl = dense_op_partial_left([c for c in b], i)[:, 0, :]
    np.testing.assert_allclose(Lb[i], np.einsum('mp,ma->pa', Bl, Xl), rtol=1e-11, atol=1e-11)
    # micro matrix = X_{!=i}^T A X_{!=i}, rows (a',m,b'), cols (a,n,b)
    Ad = tt_ops.matricize(A)
    nl, ni = Xl.shape[0], dims[i]
    nr = Xr.shape[1]
    Xne = np.einsum('la,nk,br->lnrakb', Xl, np.eye(ni), Xr).reshape(nl * ni * nr, -1)
    Md = Xne.T @ Ad @ Xne
    M = sle_als.micro_matrix(L[i], A[i], R[i])
    np.testing.assert_allclose(M, Md, rtol=1e-10, atol=1e-10)
    f = sle_als.micro_rhs(Lb[i], b[i], Rb[i]).ravel()
    np.testing.assert_allclose(f, Xne.T @ tt_ops.matricize(b), rtol=1e-10, atol=1e-10)
    v = rng.standard_normal(M.shape[1])
    np.testing.assert_allclose(sle_als.micro_matvec(L[i], A[i], R[i], v), M @ v, rtol=1e-11, atol=1e-11)


def test_identity_operator_orthonormal_stacks_are_identity():
    rng = np.random.default_rng(5)
    x = tt_ops.ortho_left(rand_tt(rng, [3, 3, 3], 3))
    I = tt_ops.eye([3, 3, 3])
    L = np.ones((1, 1, 1))
    for j in range(2):
        L = sle_als.stack_left_op(L, x[j], I[j])
        np.testing.assert_allclose(L[:, 0, :], np.eye(L.shape[0]), atol=1e-13)


def test_planted_transposed_micro_matrix_fails():
    """A micro matrix assembled with rows and columns swapped fails the dense check."""
    rng = np.random.default_rng(6)
    dims = [2, 3, 2]
    A = rand_tt(rng, dims, [1, 2, 2, 1], cols=dims)
    x = rand_tt(rng, dims, [1, 2, 2, 1])
    L = sle_als.stack_left_op(np.ones((1, 1, 1)), x[0], A[0])
    R = sle_als.stack_right_op(np.ones((1, 1, 1)), x[2], A[2])
    M = sle_als.micro_matrix(L, A[1], R)
    Xl, Xr = dense_interface_left(x, 1), dense_interface_right(x, 1)
    Xne = np.einsum('la,nk,br->lnrakb', Xl, np.eye(3), Xr).reshape(2 * 3 * 2, -1)
    Md = Xne.T @ tt_ops.matricize(A) @ Xne
    assert np.allclose(M, Md)
    assert not np.allclose(M.T, Md)  # the non-symmetric operator detects a transpose


# --------------------------------------------------------------------------
# whole solvers: golden vectors from the reference and closed forms


def test_sle_golden_laplace(golden):
    g = golden('sle_laplace_d10')
    x = sle_als.als(g['tt']('A'), g['tt']('x0'), g['tt']('b'), repeats=4, solver='solve')
    xr = g['tt']('x')
    assert tt_ops.ranks_of(x) == list(g['ranks'])
    # same LAPACK path: iterates agree to rounding (also core by core)
    diff = tt_ops.norm(tt_ops.sub(x, xr)) / tt_ops.norm(xr)
    assert diff < 1e-12, diff
    for c, cr in zip(x, xr):
        np.testing.assert_allclose(c, cr, rtol=1e-8, atol=1e-10)
    rel = tt_ops.norm(tt_ops.sub(tt_ops.dot(g['tt']('A'), x), g['tt']('b'))) / tt_ops.norm(g['tt']('b'))
    assert np.isclose(rel, float(g['relres']), rtol=1e-6)


def test_sle_golden_toeplitz_lu(golden):
    g = golden('sle_toeplitz_d6')
    for solver in ('lu', 'solve'):
        x = sle_als.als(g['tt']('A'), g['tt']('x0'), g['tt']('b'), repeats=2, solver=solver)
        assert tt_ops.ranks_of(x) == list(g['ranks'])
        xv = tt_ops.matricize(x)
        np.testing.assert_allclose(xv, tt_ops.matricize(g['tt']('x')), rtol=1e-9, atol=1e-12)
    # the reference test's own criterion (tests/test_sle.py:48-59)
    ds = g['dense_solution']
    assert np.linalg.norm(xv - ds) / np.linalg.norm(ds) < 1e-7


@pytest.mark.parametrize('name', ['evp_tfim_d8_eig', 'evp_tfim_d8_eigh', 'evp_tfim_d12_conv'])
def test_evp_golden(golden, name):
    g = golden(name)
    lam, y, it = evp_als.als(g['tt']('A'), g['tt']('x0'), repeats=int(g['repeats']), conv_eps=float(g['conv_eps']),
                             solver=str(g['solver']), sigma=float(g['sigma']))
    assert it == int(g['iterations'])
    assert abs(lam - float(g['eigenvalue'])) <= 1e-12 * abs(float(g['eigenvalue']))
    yr = g['tt']('y' if 'y_order' in g else 'x')
    assert tt_ops.ranks_of(y) == list(g['ranks'])
    d1 = tt_ops.norm(tt_ops.sub(y, yr))
    d2 = tt_ops.norm(tt_ops.add(y, yr))
    assert min(d1, d2) < 1e-8


def test_evp_golden_full_size_c2(golden):
    """The full-size eigenproblem instance (d=32, rank 16, 6 sweeps): the
    oracle reproduces the reference's eigenvalue and iteration count, and
    the eigenvalue lies above the free-fermion ground energy."""
    g = golden('evp_tfim_d32_c2')
    lam, y, it = evp_als.als(g['tt']('A'), g['tt']('x0'), repeats=int(g['repeats']), conv_eps=float(g['conv_eps']),
                             solver=str(g['solver']), sigma=float(g['sigma']))
    ref = float(g['eigenvalue'])
    assert it == int(g['iterations'])
    assert abs(lam - ref) <= 1e-12 * abs(ref), (lam, ref)
    assert tt_ops.ranks_of(y) == list(g['ranks'])
    E0 = W.tfim_ground_energy(g['J'], g['g'])
    assert E0 - 1e-12 * abs(E0) <= ref < E0 + 1e-7 * abs(E0)


def test_tfim_closed_form_vs_dense():
    for d in (2, 4, 6, 8):
        rng = np.random.default_rng(20 + d)
        J, g = rng.uniform(0.5, 1.5, d - 1), rng.uniform(0.5, 1.5, d)
        H = tt_ops.matricize(W.tfim_operator(J, g))
        np.testing.assert_allclose(H, H.T, atol=1e-14)
        assert abs(np.linalg.eigvalsh(H)[0] - W.tfim_ground_energy(J, g)) < 1e-12 * d


def test_tfim_operator_structure():
    """Uniform couplings reproduce the reference's Ising MPO (scikit_tt/models.py:31-40) densely."""
    d, J, h = 4, 1.0, 1.2
    X = np.array([[0., 1.], [1., 0.]])
    Z = np.diag([1., -1.])
    I = np.eye(2)

    def kron_at(ops):
        m = np.ones((1, 1))
        for o in ops:
            m = np.kron(m, o)
        return m
    Hd = np.zeros((16, 16))
    for i in range(d):
        Hd -= h * kron_at([X if k == i else I for k in range(d)])
    for i in range(d - 1):
        Hd -= J * kron_at([Z if k in (i, i + 1) else I for k in range(d)])
    Ht = tt_ops.matricize(W.tfim_operator([J] * (d - 1), [h] * d))
    np.testing.assert_allclose(Ht, Hd, atol=1e-14)


def test_laplace_operator_dense():
    d, n = 3, 4
    A = tt_ops.matricize(W.laplace_operator(d, n))
    D = W.laplace_1d(n)
    I = np.eye(n)
    ref = np.kron(np.kron(D, I), I) + np.kron(np.kron(I, D), I) + np.kron(np.kron(I, I), D)
    np.testing.assert_allclose(A, ref, atol=1e-12)


def test_sle_exact_for_rank_one_separable():
    """A = D x I + I x D with an eigenvector rhs: the solution is rank 1 and ALS is exact."""
    rng = np.random.default_rng(7)
    d, n = 4, 5
    A = W.laplace_operator(d, n)
    w, V = np.linalg.eigh(W.laplace_1d(n))
    b = [V[:, 0].reshape(1, n, 1, 1) for _ in range(d)]
    x0 = rand_tt(rng, [n] * d, 2)
    x = sle_als.als(A, tt_ops.ortho_right(x0), b, repeats=1)
    exact = tt_ops.scale(b, 1.0 / (d * w[0]))
    assert tt_ops.norm(tt_ops.sub(x, exact)) < 1e-10 * tt_ops.norm(exact)


def test_sle_matches_dense_solve_and_galerkin():
    rng = np.random.default_rng(8)
    d, n = 5, 3
    A = W.laplace_operator(d, n)
    b = rand_tt(rng, [n] * d, 2)
    x0 = rand_tt(rng, [n] * d, 9)  # rank 9 = full for n^2, exact solution representable
    x = sle_als.als(A, x0, b, repeats=2)
    xd = np.linalg.solve(tt_ops.matricize(A), tt_ops.matricize(b))
    assert np.linalg.norm(tt_ops.matricize(x) - xd) / np.linalg.norm(xd) < 1e-10
    # Galerkin condition of the last local solve (core 0): X_{!=0}^T (A x - b) = 0
    x1 = sle_als.als(A, rand_tt(rng, [n] * d, 2), b, repeats=1)
    r = tt_ops.matricize(A) @ tt_ops.matricize(x1) - tt_ops.matricize(b)
    Xr = dense_interface_right(x1, 0)
    Xne = np.einsum('nk,br->nrkb', np.eye(n), Xr).reshape(n * Xr.shape[1], -1)
    assert np.linalg.norm(Xne.T @ r) < 1e-9 * np.linalg.norm(tt_ops.matricize(b))
    # and the rank-2 ALS solution is not the exact one (the check is not vacuous)
    assert np.linalg.norm(r) > 1e-6 * np.linalg.norm(tt_ops.matricize(b))


def test_sle_gauge_and_resume_invariance():
    A, x0, b = W.laplace_small(d=6, n=5, rank=3)
    x = sle_als.als(A, x0, b, repeats=3)
    xg = sle_als.als(A, tt_ops.ortho_right(x0), b, repeats=3)
    assert tt_ops.norm(tt_ops.sub(x, xg)) < 1e-11 * tt_ops.norm(x)
    xr = sle_als.als(A, sle_als.als(A, x0, b, repeats=2), b, repeats=1)
    assert tt_ops.norm(tt_ops.sub(x, xr)) < 1e-10 * tt_ops.norm(x)


def test_planted_row_col_swap_in_stack_fails():
    """Contracting x with the operator ROW index (instead of the column) fails the dense check."""
    rng = np.random.default_rng(9)
    dims = [3, 3]
    A = rand_tt(rng, dims, [1, 2, 1], cols=dims)
    x = rand_tt(rng, dims, [1, 2, 1])
    L1 = sle_als.stack_left_op(np.ones((1, 1, 1)), x[0], A[0])
    Lbad = sle_als.stack_left_op(np.ones((1, 1, 1)), x[0], A[0].transpose(0, 2, 1, 3))
    Xl = dense_interface_left(x, 1)
    Ld = np.einsum('na,mnp,mc->apc', Xl, dense_op_partial_left(A, 1), Xl)
    assert np.allclose(L1, Ld) and not np.allclose(Lbad, Ld)


def test_planted_selection_rule_fails_golden(golden, monkeypatch):
    """Picking the eigenvalue FARTHEST from sigma breaks agreement with the reference's vectors."""
    g = golden('evp_tfim_d8_eig')
    orig = evp_als.local_eigen

    def planted(M, B, k, solver, sigma, real):
        vals, vecs = sla.eig(M)
        order = np.argsort(-np.abs(vals - sigma))
        return np.real(vals[order[:k]]), vecs[:, order[:k]]
    monkeypatch.setattr(evp_als, 'local_eigen', planted)
    lam, y, it = evp_als.als(g['tt']('A'), g['tt']('x0'), repeats=int(g['repeats']), conv_eps=float(g['conv_eps']),
                             solver='eig', sigma=float(g['sigma']))
    assert abs(lam - float(g['eigenvalue'])) > 1e-3


def test_planted_dropped_rhs_stack_fails_golden(golden, monkeypatch):
    """Skipping the rhs stack update (a dropped term) breaks agreement with the reference iterate."""
    g = golden('sle_laplace_d10')
    monkeypatch.setattr(sle_als, 'stack_left_rhs', lambda prev, xc, bc: np.ones((bc.shape[3], xc.shape[3])))
    x = sle_als.als(g['tt']('A'), g['tt']('x0'), g['tt']('b'), repeats=4, solver='solve')
    xr = g['tt']('x')
    # the final core-0 solve is exact on the (slightly wrong) interfaces, so the
    # damage is small -- but far above the 1e-12 agreement test_sle_golden_laplace asks
    assert tt_ops.norm(tt_ops.sub(x, xr)) / tt_ops.norm(xr) > 1e-9


def test_evp_closed_form_small():
    H, y0, J, g = W.tfim_chain(seed=1, d=6, rank=8)
    y0 = tt_ops.ortho_right(y0)
    y0 = tt_ops.scale(y0, 1.0 / tt_ops.norm(y0))
    lam, y, it = evp_als.als(H, y0, repeats=4, conv_eps=0, solver='eig', sigma=-100)
    assert abs(lam - W.tfim_ground_energy(J, g)) < 1e-10
    assert it == 4
    # returned TT is a unit eigenvector: Rayleigh quotient equals lambda
    Hd = tt_ops.matricize(H)
    v = tt_ops.matricize(y)
    assert abs(v @ Hd @ v - lam) < 1e-10 and abs(np.linalg.norm(v) - 1) < 1e-12


def test_evp_options_fixture_against_dense():
    """The reference-written evp-options fixture against dense eigenvalues of
    its own stored operators: the Ulam test problem within the reference
    test's tolerance (5e-3), the generalised TFIM problem and the deflated
    (first excited) one to the ALS accuracy at these ranks."""
    import scipy.linalg as sla
    from tests.conftest import load_golden
    g = load_golden('evp_options')
    A = tt_ops.matricize(g['tt']('ulam_A'))
    top3 = np.linalg.eigvalsh(0.5 * (A + A.T))[::-1][:3]
    np.testing.assert_allclose(top3, g['ulam_dense_top3'], rtol=1e-12)
    assert np.all(np.abs(g['ulam_vals'] - top3) <= 5e-3 * np.abs(top3))
    A = tt_ops.matricize(g['tt']('ulam_ns_A'))
    ev = np.linalg.eigvals(A)
    top = np.real(ev[np.argsort(-np.real(ev))][:3])
    np.testing.assert_allclose(top, g['ulam_ns_dense_top3'], rtol=1e-10)
    assert abs(float(g['ulam_ns_val']) - 1.0) <= 5e-3
    assert np.all(np.abs(g['ulam_ns3_vals'] - top) <= 5e-3 * np.abs(top))
    H = tt_ops.matricize(g['tt']('tfim_A'))
    M = tt_ops.matricize(g['tt']('tfim_M'))
    assert np.all(np.linalg.eigvalsh(M) > 0)
    lowest = sla.eigh(H, M, eigvals_only=True)[0]
    assert abs(float(g['gevp_val']) - lowest) <= 1e-6 * abs(lowest)
    two = np.linalg.eigvalsh(H)[:2]
    assert abs(float(g['defl_val']) - two[1]) <= 1e-6 * abs(two[1])
    assert abs(float(g['defl_prev_val']) - two[0]) <= 1e-6 * abs(two[0])
