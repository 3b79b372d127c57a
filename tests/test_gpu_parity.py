"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Tolerances: pieces (stack contractions, matvecs, micro matrices) 1e-12
relative; local Krylov solves to their stated tolerance; whole-solver
iterates 1e-7 relative (the reference tests' own bar, tests/test_sle.py:17)
and eigenvalues 1e-10 relative.  TT ranks must be identical.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import tt_ops, sle_als, evp_als
from scikit_tt_amd import workloads as W
from tests.conftest import load_golden

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def env():
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    from scikit_tt_amd import _lib
    torch.cuda.set_device(0)
    lib = _lib.load()
    return lib, _lib.context(0), _lib


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def empty(*shape):
    return torch.empty(shape, dtype=torch.float64, device='cuda')


@pytest.fixture(params=[(61, 0), (61, 1), (0, 0)], ids=['onchip', 'onchip3', 'global'])
def arena(request, env):
    """Run a piece on the engine's on-chip (LDS, matrix-core) path where the
    problem fits -- two-stage matvec (operator core merged with the right
    stack) or the three-stage one -- or force the global (HBM/L2) path."""
    lib, ctx, _ = env
    kb, variant = request.param
    assert lib.ttals_set_arena_kb(ctx, kb) == 0
    assert lib.ttals_set_matvec_variant(ctx, variant) == 0
    yield kb
    assert lib.ttals_set_arena_kb(ctx, 61) == 0
    assert lib.ttals_set_matvec_variant(ctx, 0) == 0


SHAPES = [(1, 1, 2, 3, 2), (3, 2, 4, 3, 5), (5, 2, 8, 2, 5), (20, 3, 8, 2, 17), (16, 3, 2, 3, 16), (2, 4, 16, 4, 48)]


@pytest.mark.parametrize('a,P,n,Q,b', SHAPES)
def test_stacks_matvec_micro(env, arena, a, P, n, Q, b):
    lib, ctx, _ = env
    rng = np.random.default_rng(a * 100 + b)
    L = rng.standard_normal((a, P, a))
    R = rng.standard_normal((b, Q, b))
    Ac = rng.standard_normal((P, n, n, Q))
    x = rng.standard_normal((a, n, 1, b))
    dL, dR, dA, dx = dev(L), dev(R), dev(Ac), dev(x)
    out = empty(b, Q, b)
    assert lib.ttals_stack_left_op(ctx, None, a, P, n, Q, b, ptr(dL), ptr(dA), ptr(dx), ptr(out)) == 0
    ref = sle_als.stack_left_op(L, x, Ac)
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    out = empty(a, P, a)
    assert lib.ttals_stack_right_op(ctx, None, a, P, n, Q, b, ptr(dR), ptr(dA), ptr(dx), ptr(out)) == 0
    ref = sle_als.stack_right_op(R, x, Ac)
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # rhs stacks with p = P, q = Q ranks
    Lb = rng.standard_normal((P, a))
    Rb = rng.standard_normal((Q, b))
    Bc = rng.standard_normal((P, n, 1, Q))
    dLb, dRb, dB = dev(Lb), dev(Rb), dev(Bc)
    out = empty(Q, b)
    assert lib.ttals_stack_left_rhs(ctx, None, a, P, n, Q, b, ptr(dLb), ptr(dB), ptr(dx), ptr(out)) == 0
    np.testing.assert_allclose(out.cpu().numpy(), sle_als.stack_left_rhs(Lb, x, Bc), rtol=1e-12, atol=1e-11)
    out = empty(P, a)
    assert lib.ttals_stack_right_rhs(ctx, None, a, P, n, Q, b, ptr(dRb), ptr(dB), ptr(dx), ptr(out)) == 0
    np.testing.assert_allclose(out.cpu().numpy(), sle_als.stack_right_rhs(Rb, x, Bc), rtol=1e-12, atol=1e-11)
    out = empty(a * n * b)
    assert lib.ttals_micro_rhs(ctx, None, a, P, n, Q, b, ptr(dLb), ptr(dB), ptr(dRb), ptr(out)) == 0
    np.testing.assert_allclose(out.cpu().numpy(), sle_als.micro_rhs(Lb, Bc, Rb).ravel(), rtol=1e-12, atol=1e-11)
    # matvec and dense micro matrix
    N = a * n * b
    v = rng.standard_normal(N)
    y = empty(N)
    dv = dev(v)
    assert lib.ttals_micro_matvec(ctx, None, a, P, n, Q, b, ptr(dL), ptr(dA), ptr(dR), ptr(dv), ptr(y)) == 0
    ref = sle_als.micro_matvec(L, Ac, R, v)
    np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())
    if N <= 600:
        M = empty(N, N)
        assert lib.ttals_micro_matrix(ctx, None, a, P, n, Q, b, ptr(dL), ptr(dA), ptr(dR), ptr(M)) == 0
        Mr = sle_als.micro_matrix(L, Ac, R)
        np.testing.assert_allclose(M.cpu().numpy(), Mr, rtol=1e-12, atol=1e-12 * np.abs(Mr).max())


def spd_local(rng, a, P, n, Q, b, shift=5.0):
    """Random SPD local operator: symmetric slices, PSD stacks, identity shift in slice 0."""
    Ac = rng.standard_normal((P, n, n, Q))
    Ac = Ac + Ac.transpose(0, 2, 1, 3)
    Ac[0, :, :, 0] += shift * n * np.eye(n) * 4
    def psd(r, k):
        S = rng.standard_normal((r, k, r)) * 0.1
        S = S + S.transpose(2, 1, 0)
        S[:, 0, :] += np.eye(r) * 3
        return S
    return psd(a, P), Ac, psd(b, Q)


@pytest.mark.parametrize('a,P,n,Q,b,method', [(2, 2, 3, 2, 4, 1), (4, 2, 6, 2, 4, 1), (5, 2, 8, 2, 5, 2),
                                              (8, 3, 4, 3, 8, 2), (5, 2, 8, 2, 5, 3), (8, 3, 4, 3, 9, 3)])
def test_local_solve(env, arena, a, P, n, Q, b, method):
    lib, ctx, _ = env
    rng = np.random.default_rng(7 + method)
    L, Ac, R = spd_local(rng, a, P, n, Q, b)
    if method == 3:  # general (non-symmetric) operator for GMRES
        Ac = Ac + 0.3 * rng.standard_normal(Ac.shape)
    M = sle_als.micro_matrix(L, Ac, R)
    N = M.shape[0]
    f = rng.standard_normal(N)
    v = torch.zeros(N, dtype=torch.float64, device='cuda')
    its = ctypes.c_int()
    keep = [dev(L), dev(Ac), dev(R), dev(f)]  # device copies must outlive the call
    rc = lib.ttals_local_solve(ctx, None, a, P, n, Q, b, *[ptr(t) for t in keep], ptr(v),
                               method, 1 if method == 2 else 0, 1e-12, 4000, ctypes.byref(its))
    assert rc == 0
    x = v.cpu().numpy()
    assert np.linalg.norm(M @ x - f) <= 1e-10 * np.linalg.norm(f)
    np.testing.assert_allclose(x, np.linalg.solve(M, f), rtol=1e-8, atol=1e-9 * np.abs(x).max())


@pytest.mark.parametrize('eps', [1e-14, 1e-7])
def test_local_gmres_short_run_true_residual(env, arena, eps):
    """GMRES runs of one or two steps return without a confirming matvec: the
    least-squares residual they accept must be the true residual.  M = I +
    eps (I x E x I) converges in one step (eps = 1e-14) or two (1e-7); the
    true residual of the returned x is checked on the host against the
    solver's acceptance bound (10 tol)."""
    lib, ctx, _ = env
    a, P, n, Q, b = 8, 1, 8, 1, 8
    rng = np.random.default_rng(3)
    L = np.eye(a)[:, None, :].copy()
    R = np.eye(b)[:, None, :].copy()
    Ac = (np.eye(n) + eps * rng.standard_normal((n, n)))[None, :, :, None].copy()
    M = sle_als.micro_matrix(L, Ac, R)
    N = M.shape[0]
    f = rng.standard_normal(N)
    v = torch.zeros(N, dtype=torch.float64, device='cuda')
    its = ctypes.c_int()
    keep = [dev(L), dev(Ac), dev(R), dev(f)]
    rc = lib.ttals_local_solve(ctx, None, a, P, n, Q, b, *[ptr(t) for t in keep], ptr(v), 3, 0, 1e-12, 4000,
                               ctypes.byref(its))
    assert rc == 0
    assert 1 <= its.value <= 2, its.value
    x = v.cpu().numpy()
    assert np.linalg.norm(M @ x - f) <= 10 * 1e-12 * np.linalg.norm(f)


def test_local_lu_singular_detected(env):
    lib, ctx, _lib = env
    a, P, n, Q, b = 2, 1, 3, 1, 2
    Ac = np.zeros((P, n, n, Q))
    keep = [dev(np.ones((a, P, a))), dev(Ac), dev(np.ones((b, Q, b))), dev(np.ones(a * n * b)),
            torch.zeros(a * n * b, dtype=torch.float64, device='cuda')]
    rc = lib.ttals_local_solve(ctx, None, a, P, n, Q, b, *[ptr(t) for t in keep], 1, 0, 1e-12, 10, None)
    assert rc == _lib.S_SINGULAR


@pytest.mark.parametrize('N_shape,mode', [((2, 2, 3, 2, 4), 0), ((4, 2, 6, 2, 4), 1), ((8, 3, 4, 3, 8), 0),
                                          ((8, 3, 4, 3, 8), 2), ((2, 3, 2, 3, 4), 2)])
def test_local_eig(env, arena, N_shape, mode):
    lib, ctx, _ = env
    a, P, n, Q, b = N_shape
    rng = np.random.default_rng(11)
    L, Ac, R = spd_local(rng, a, P, n, Q, b, shift=0.0)
    M = sle_als.micro_matrix(L, Ac, R)
    w = np.linalg.eigvalsh(M)
    sigma = {0: 0.0, 1: 0.0, 2: w[0] - 100.0}[mode]
    v = dev(rng.standard_normal(M.shape[0]))
    lam, its = ctypes.c_double(), ctypes.c_int()
    keep = [dev(L), dev(Ac), dev(R)]
    rc = lib.ttals_local_eig(ctx, None, a, P, n, Q, b, *[ptr(t) for t in keep], ptr(v), mode, sigma,
                             1e-12, 4000, ctypes.byref(lam), ctypes.byref(its))
    assert rc == 0
    target = w[-1] if mode == 1 else w[0]
    assert abs(lam.value - target) <= 1e-10 * np.abs(w).max()
    x = v.cpu().numpy()
    assert abs(np.linalg.norm(x) - 1) < 1e-12
    assert np.linalg.norm(M @ x - lam.value * x) <= 1e-8 * np.abs(w).max()


def test_local_eig_restarts(env, arena):
    """A lowest eigenvalue with a small relative gap needs several Lanczos
    restart cycles (63 steps each); the restarted solve converges to it.
    The local operator is a Kronecker product of diagonal stacks and core,
    so its spectrum is known exactly."""
    lib, ctx, _ = env
    a, P, n, Q, b = 4, 1, 8, 1, 16
    L = np.zeros((a, 1, a))
    L[:, 0, :] = np.diag([1.0, 1.3, 1.7, 2.0])
    Ac = np.zeros((1, n, n, 1))
    Ac[0, :, :, 0] = np.diag(np.linspace(1.0, 3.0, n))
    R = np.zeros((b, 1, b))
    R[:, 0, :] = np.diag(np.concatenate(([1.0, 1.05], np.linspace(1.2, 4.0, b - 2))))
    M = sle_als.micro_matrix(L, Ac, R)
    w = np.linalg.eigvalsh(M)
    v = dev(np.random.default_rng(3).standard_normal(M.shape[0]))
    lam, its = ctypes.c_double(), ctypes.c_int()
    keep = [dev(L), dev(Ac), dev(R)]
    rc = lib.ttals_local_eig(ctx, None, a, P, n, Q, b, *[ptr(t) for t in keep], ptr(v), 0, 0.0,
                             1e-12, 4000, ctypes.byref(lam), ctypes.byref(its))
    assert rc == 0
    assert its.value > 63  # at least one restart
    assert abs(lam.value - w[0]) <= 1e-10 * np.abs(w).max()
    x = v.cpu().numpy()
    assert abs(np.linalg.norm(x) - 1) < 1e-12
    assert np.linalg.norm(M @ x - lam.value * x) <= 1e-8 * np.abs(w).max()


def test_local_eig_nearest_interior_small(env):
    """Local problem inside the Krylov space: the eigenvalue nearest an interior sigma (evp.py:429)."""
    lib, ctx, _ = env
    a, P, n, Q, b = 2, 2, 3, 2, 3
    rng = np.random.default_rng(12)
    L, Ac, R = spd_local(rng, a, P, n, Q, b, shift=0.0)
    M = sle_als.micro_matrix(L, Ac, R)
    w = np.linalg.eigvalsh(M)
    sigma = 0.5 * (w[7] + w[8]) + 0.1 * (w[8] - w[7])
    v = dev(rng.standard_normal(M.shape[0]))
    lam = ctypes.c_double()
    keep = [dev(L), dev(Ac), dev(R)]
    assert lib.ttals_local_eig(ctx, None, a, P, n, Q, b, *[ptr(t) for t in keep], ptr(v), 2, sigma,
                               1e-12, 4000, ctypes.byref(lam), None) == 0
    assert abs(lam.value - w[np.argmin(np.abs(w - sigma))]) < 1e-9 * np.abs(w).max()


@pytest.mark.parametrize('m,c', [(12, 5), (5, 12), (200, 48), (7, 7), (1, 9), (9, 1), (32, 16), (64, 64),
                                 (64, 40), (30, 64), (65, 16), (200, 80), (130, 130), (100, 300), (520, 70)])
def test_qr(env, m, c):
    """Small panels (m, c <= 64) take the DPP-row QR; larger tall ones
    (m >= c) the blocked Householder QR; wide ones the column sweep."""
    lib, ctx, _ = env
    Y = np.random.default_rng(m * c).standard_normal((m, c))
    k = min(m, c)
    q, r = empty(m, k), empty(k, c)
    dY = dev(Y)
    assert lib.ttals_qr(ctx, None, m, c, ptr(dY), ptr(q), ptr(r)) == 0
    Q, Rr = q.cpu().numpy(), r.cpu().numpy()
    np.testing.assert_allclose(Q @ Rr, Y, atol=1e-13 * np.abs(Y).max() * max(m, c))
    np.testing.assert_allclose(Q.T @ Q, np.eye(k), atol=1e-13 * max(m, c))
    assert np.allclose(np.tril(Rr, -1), 0.0)
    # same factorisation as LAPACK up to the signs of R's rows
    Qr, Rl = sla.qr(Y, mode='economic')
    np.testing.assert_allclose(np.abs(Rr), np.abs(Rl), rtol=1e-10, atol=1e-12 * np.abs(Y).max())


@pytest.mark.parametrize('m,c,rank,cond', [(300, 40, 20, 1.0), (520, 48, 48, 1e12), (768, 48, 47, 1.0)])
def test_qr_ill_conditioned_and_rank_deficient(env, m, c, rank, cond):
    """Tall panels of numerical rank below c or of condition number 1e12
    (blocked Householder): Q R = Y and Q orthonormal to working accuracy; for
    the full-rank panel R matches LAPACK's up to row signs."""
    lib, ctx, _ = env
    rng = np.random.default_rng(m + c + rank)
    U, _ = np.linalg.qr(rng.standard_normal((m, rank)))
    V, _ = np.linalg.qr(rng.standard_normal((c, rank)))
    Y = (U * np.geomspace(1.0, 1.0 / cond, rank)) @ V.T
    q, r = empty(m, c), empty(c, c)
    dY = dev(Y)
    assert lib.ttals_qr(ctx, None, m, c, ptr(dY), ptr(q), ptr(r)) == 0
    Q, Rr = q.cpu().numpy(), r.cpu().numpy()
    np.testing.assert_allclose(Q @ Rr, Y, atol=1e-13 * np.abs(Y).max() * m)
    np.testing.assert_allclose(Q.T @ Q, np.eye(c), atol=1e-12 * m)
    assert np.allclose(np.tril(Rr, -1), 0.0)
    if rank == c:
        _, Rl = sla.qr(Y, mode='economic')
        np.testing.assert_allclose(np.abs(Rr), np.abs(Rl), atol=1e-10 * np.abs(Rl).max())


# --------------------------------------------------------------------------
# TT algebra through the TT class


def test_tt_algebra(env):
    from scikit_tt_amd.tensor_train import TT
    import scikit_tt_amd.tensor_train as tt
    rng = np.random.default_rng(5)
    A = W.random_tt(rng, [2, 3, 2, 3], [1, 2, 3, 2, 1], col_dims=[3, 2, 2, 3])
    x = W.random_tt(rng, [3, 2, 2, 3], [1, 3, 4, 2, 1])
    y = W.random_tt(rng, [3, 2, 2, 3], [1, 2, 2, 3, 1])
    tA, tx, ty = TT(A), TT(x), TT(y)
    np.testing.assert_allclose(tA.dot(tx).matricize(), tt_ops.matricize(tt_ops.dot(A, x)), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose((tA @ tx).matricize(), tt_ops.matricize(A) @ tt_ops.matricize(x), rtol=1e-11,
                               atol=1e-11)
    np.testing.assert_allclose((tx + ty).matricize(), tt_ops.matricize(tt_ops.add(x, y)), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose((tx - ty).matricize(), tt_ops.matricize(tt_ops.sub(x, y)), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose((2.5 * tx).matricize(), 2.5 * tt_ops.matricize(x), rtol=1e-13)
    np.testing.assert_allclose((tx * -1.5).matricize(), -1.5 * tt_ops.matricize(x), rtol=1e-13)
    for c, cr in zip((tx + ty).cores, tt_ops.add(x, y)):
        np.testing.assert_array_equal(c, cr)  # the block layout itself
    assert abs(tx.norm() - tt_ops.norm(x)) <= 1e-13 * tt_ops.norm(x)
    assert abs(tA.norm() - tt_ops.norm(A)) <= 1e-13 * tt_ops.norm(A)
    xr = tt_ops.ortho_right(x)
    t = TT(x).ortho_right()
    assert t.ranks == tt_ops.ranks_of(xr)
    np.testing.assert_allclose(t.matricize(), tt_ops.matricize(x), rtol=1e-12, atol=1e-12)
    for c in t.cores[1:]:
        m = c.reshape(c.shape[0], -1)
        np.testing.assert_allclose(m @ m.T, np.eye(m.shape[0]), atol=1e-13)
    t = TT(x).ortho_left()
    assert t.ranks == tt_ops.ranks_of(tt_ops.ortho_left(x))
    np.testing.assert_allclose(t.matricize(), tt_ops.matricize(x), rtol=1e-12, atol=1e-12)
    u = TT(W.random_tt(rng, [1, 1], [1, 2, 1]))
    assert isinstance(u.dot(TT(W.random_tt(rng, [1, 1], [1, 2, 1]))), float)
    with pytest.raises(ValueError):
        tx.dot(tx)
    with pytest.raises(TypeError):
        tx + 1
    with pytest.raises(ValueError):
        TT([np.zeros((1, 2, 1, 2)), np.zeros((3, 2, 1, 1))])
    with pytest.raises(TypeError):
        TT('cores')
    assert tt.eye([2, 3]).matricize().shape == (6, 6)
    # ranks [1, 2, 1]: every entry is the sum over the 2 rank paths of 1 * 1
    np.testing.assert_allclose(tt.ones([2, 3], [1, 1], ranks=2).matricize(), 2 * np.ones(6))


# --------------------------------------------------------------------------
# whole solvers


def _tt(cores):
    from scikit_tt_amd.tensor_train import TT
    return TT([np.asarray(c) for c in cores])


def test_sle_golden_laplace(env):
    import scikit_tt_amd.solvers.sle as sle
    g = load_golden('sle_laplace_d10')
    A, x0, b = g['tt']('A'), g['tt']('x0'), g['tt']('b')
    (x,), (info,) = sle.als_batch([_tt(A)], [_tt(x0)], [_tt(b)], repeats=4, return_info=True)
    assert info.status == 0 and info.unconverged == 0, info
    assert x.ranks == list(g['ranks'])
    xr = g['tt']('x')
    diff = tt_ops.norm(tt_ops.sub(x.cores[:], xr)) / tt_ops.norm(xr)
    assert diff < 1e-7, diff
    rel = (_tt(A).dot(x) - _tt(b)).norm() / _tt(b).norm()
    assert abs(rel - float(g['relres'])) < 1e-3 * float(g['relres'])


@pytest.mark.parametrize('solver', ['lu', 'solve'])
def test_sle_golden_toeplitz(env, solver):
    """The reference test's indefinite Toeplitz pattern (tests/test_sle.py:20-59), smaller."""
    import scikit_tt_amd.solvers.sle as sle
    g = load_golden('sle_toeplitz_d6')
    x = sle.als(_tt(g['tt']('A')), _tt(g['tt']('x0')), _tt(g['tt']('b')), repeats=2, solver=solver)
    assert x.ranks == list(g['ranks'])
    np.testing.assert_allclose(x.matricize(), tt_ops.matricize(g['tt']('x')), rtol=1e-9, atol=1e-12)
    ds = g['dense_solution']
    assert np.linalg.norm(x.matricize() - ds) / np.linalg.norm(ds) < 1e-7


@pytest.mark.parametrize('name', ['evp_tfim_d8_eig', 'evp_tfim_d8_eigh', 'evp_tfim_d12_conv'])
def test_evp_golden(env, name):
    import scikit_tt_amd.solvers.evp as evp
    g = load_golden(name)
    lam, y, it = evp.als(_tt(g['tt']('A')), _tt(g['tt']('x0')), repeats=int(g['repeats']),
                         conv_eps=float(g['conv_eps']), solver=str(g['solver']), sigma=float(g['sigma']))
    assert it == int(g['iterations'])
    assert abs(lam - float(g['eigenvalue'])) <= 1e-10 * abs(float(g['eigenvalue']))
    assert y.ranks == list(g['ranks'])
    yr = g['tt']('x')
    d1 = tt_ops.norm(tt_ops.sub(y.cores[:], yr))
    d2 = tt_ops.norm(tt_ops.add(y.cores[:], yr))
    assert min(d1, d2) < 1e-6


def test_sle_laplace_small_vs_oracle(env):
    import scikit_tt_amd.solvers.sle as sle
    A, x0, b = W.laplace_small()
    for reps in (1, 2):
        x = sle.als(_tt(A), _tt(x0), _tt(b), repeats=reps)
        xr = sle_als.als(A, x0, b, repeats=reps)
        assert x.ranks == tt_ops.ranks_of(xr)
        assert tt_ops.norm(tt_ops.sub(x.cores[:], xr)) / tt_ops.norm(xr) < 1e-7


def _short_train(d):
    """A symmetric positive definite operator of order d, x0 and b (order 1:
    one dense 5x5 core; orders 2, 3: the Laplacian with n=4, rank 2)."""
    if d > 1:
        return W.laplace_small(d=d, n=4, rank=2)
    rng = np.random.default_rng(3)
    M = rng.standard_normal((5, 5))
    M = M @ M.T + 5.0 * np.eye(5)
    return [M.reshape(1, 5, 5, 1)], [rng.standard_normal((1, 5, 1, 1))], [np.ones((1, 5, 1, 1))]


@pytest.mark.parametrize('d', [1, 2, 3])
def test_short_trains_vs_oracle(env, d):
    """Orders 1-3 (one core: no forward step, one local solve per sweep):
    sle and evp ('eig' nearest sigma, 'eigh' largest) against the oracle."""
    import scikit_tt_amd.solvers.sle as sle
    import scikit_tt_amd.solvers.evp as evp
    A, x0, b = _short_train(d)
    for reps in (1, 2):
        x = sle.als(_tt(A), _tt(x0), _tt(b), repeats=reps)
        xr = sle_als.als(A, x0, b, repeats=reps)
        assert x.ranks == tt_ops.ranks_of(xr)
        # compared as full tensors: for one core, x - y of the reference writes
        # both operands into the same 1x1 rank block (tensor_train.py:313-333)
        # and returns -y, which the device addition reproduces
        xd, xrd = x.matricize().ravel(), tt_ops.matricize(xr).ravel()
        assert np.linalg.norm(xd - xrd) / np.linalg.norm(xrd) < 1e-10
        np.testing.assert_allclose((x - _tt(xr)).matricize().ravel(),
                                   tt_ops.matricize(tt_ops.sub(x.cores[:], xr)).ravel(), rtol=1e-14, atol=1e-15)
    dense = tt_ops.matricize(A)
    for solver, sigma in (('eig', -100.0), ('eigh', 1.0)):
        lam, y, it = evp.als(_tt(A), _tt(x0), repeats=3, conv_eps=0, solver=solver, sigma=sigma)
        lr, yr, itr = evp_als.als(A, x0, repeats=3, conv_eps=0, solver=solver, sigma=sigma)
        assert it == itr
        assert abs(lam - lr) < 1e-11 * abs(lr), (solver, lam, lr)
        # eigenvectors agree up to sign
        ov = abs(float(y.matricize().ravel() @ tt_ops.matricize(yr).ravel()))
        assert abs(ov - 1.0) < 1e-9, (solver, ov)
        if d == 1:  # one core: the exact extreme eigenvalue
            ev = np.linalg.eigvalsh(dense)
            assert abs(lam - (ev[0] if solver == 'eig' else ev[-1])) < 1e-12 * abs(lam)


def test_evp_tfim_chain_full_size(env):
    """TFIM chain d=32, rank 16, 6 sweeps: lowest eigenvalue vs the free-fermion closed form."""
    import scikit_tt_amd.solvers.evp as evp
    H, y0, J, g = W.tfim_chain(seed=0, d=32, rank=16)
    y0 = tt_ops.ortho_right(y0)
    y0 = tt_ops.scale(y0, 1.0 / tt_ops.norm(y0))
    lam, y, it = evp.als(_tt(H), _tt(y0), repeats=6, conv_eps=0, solver='eig', sigma=-100)
    E0 = W.tfim_ground_energy(J, g)
    assert it == 6
    assert lam >= E0 - 1e-9 * abs(E0)          # variational
    assert lam - E0 < 1e-7 * abs(E0), (lam, E0)
    assert abs(y.norm() - 1.0) < 1e-10
    # the Rayleigh quotient of the returned (unit) TT equals the returned eigenvalue
    rq = y.transpose().dot(_tt(H).dot(y))
    assert abs(rq - lam) < 1e-10 * abs(lam)


def test_batch_equals_alone(env):
    """Each system's result is the same alone and in a mixed batch of uneven orders (one workgroup per system)."""
    import scikit_tt_amd.solvers.sle as sle
    systems = [W.shifted_batch_instance(s, d_range=(4, 8), n_choices=(2, 4), r_range=(2, 6)) for s in range(6)]
    systems.append(W.laplace_small(d=6, n=5, rank=3))
    systems += [_short_train(1), _short_train(2)]  # ragged orders down to one core
    As = [_tt(A) for A, _, _ in systems]
    Xs = [_tt(x) for _, x, _ in systems]
    Bs = [_tt(b) for _, _, b in systems]
    together, infos = sle.als_batch(As, Xs, Bs, repeats=2, return_info=True)
    for k in range(len(systems)):
        (alone,), _ = sle.als_batch([As[k]], [Xs[k]], [Bs[k]], repeats=2, return_info=True)
        assert alone.ranks == together[k].ranks
        for c1, c2 in zip(alone.cores, together[k].cores):
            np.testing.assert_array_equal(c1, c2)


def test_sle_errors(env):
    import scikit_tt_amd.solvers.sle as sle
    import scikit_tt_amd.solvers.evp as evp
    A, x0, b = W.laplace_small(d=4, n=3, rank=2)
    with pytest.raises(NotImplementedError):
        sle.als(_tt(A), _tt(x0), _tt(b), solver='gmres')
    with pytest.raises(TypeError):
        sle.als(A, _tt(x0), _tt(b))
    with pytest.raises(ValueError):
        sle.als(_tt(A), _tt(x0[:3] + [np.ones((2, 3, 1, 1))]), _tt(b[:3] + [np.ones((1, 2, 1, 1))]))
    zero = [np.zeros_like(c) for c in A]
    with pytest.raises(np.linalg.LinAlgError):
        sle.als(_tt(zero), _tt(x0), _tt(b))
    # (number_ev > 1 and operator_gevp run: tests/test_gpu_evp_options.py)
    with pytest.raises(NotImplementedError):
        evp.als(_tt(A), _tt(x0), solver='eigs')
    with pytest.raises(NotImplementedError):
        evp.als(_tt(A), _tt(x0), real=False)


def _gapped_tt(rng, dims, rank, decay):
    """A TT whose bond spectra have clear gaps: core_i = Q_i diag(decay) ..."""
    cores = []
    r0 = 1
    for i, n in enumerate(dims):
        r1 = 1 if i == len(dims) - 1 else rank
        c = rng.standard_normal((r0, n, 1, r1))
        if i < len(dims) - 1:
            c = c * np.asarray(decay)[None, None, None, :r1]
        cores.append(c)
        r0 = r1
    return cores


@pytest.mark.parametrize('thr,maxr', [(1e-6, np.inf), (1e-3, np.inf), (1e-9, 3), (0, 2), (1e-6, [1, 4, 3, 5, 2, 1])])
def test_truncated_ortho_ranks_and_tensor(env, thr, maxr):
    """ortho with threshold/max_rank: identical ranks to the reference rule, same truncated tensor."""
    from scikit_tt_amd.tensor_train import TT
    rng = np.random.default_rng(21)
    decay = [1.0, 0.3, 1e-2, 1e-5, 1e-8, 1e-12]
    x = _gapped_tt(rng, [3, 4, 3, 4, 3], 6, decay)
    ref = tt_ops.ortho(x, threshold=thr, max_rank=maxr)
    t = TT(x).ortho(threshold=thr, max_rank=maxr)
    assert t.ranks == tt_ops.ranks_of(ref)
    np.testing.assert_allclose(t.matricize(), tt_ops.matricize(ref), rtol=1e-9,
                               atol=1e-9 * np.abs(tt_ops.matricize(ref)).max())
    for c in t.cores[1:]:
        m = c.reshape(c.shape[0], -1)
        np.testing.assert_allclose(m @ m.T, np.eye(m.shape[0]), atol=1e-12)
    # the constructor's truncation path (tensor_train.py:190-192)
    t2 = TT(x, threshold=thr, max_rank=maxr)
    assert t2.ranks == t.ranks


def test_ortho_left_truncated(env):
    from scikit_tt_amd.tensor_train import TT
    rng = np.random.default_rng(22)
    x = _gapped_tt(rng, [2, 5, 3, 4], 5, [1.0, 1e-1, 1e-4, 1e-9, 1e-14])
    ref = tt_ops.ortho_left(x, threshold=1e-7)
    t = TT(x).ortho_left(threshold=1e-7)
    assert t.ranks == tt_ops.ranks_of(ref)
    np.testing.assert_allclose(t.matricize(), tt_ops.matricize(ref), rtol=1e-8, atol=1e-10)


def test_from_host_batch_and_norm_leave_cores_untouched(env):
    from scikit_tt_amd.tensor_train import TT, from_host_batch
    rng = np.random.default_rng(21)
    lists = [W.random_tt(rng, [2, 3, 2, 2, 3], [1, 2, 4, 3, 2, 1]) for _ in range(3)] + \
            [W.random_tt(rng, [4], [1, 1])]
    tts = from_host_batch(lists)
    for cl, t in zip(lists, tts):
        assert t.ranks == [c.shape[0] for c in cl] + [1]
        for c, dc in zip(cl, t.device_cores):
            np.testing.assert_array_equal(dc.cpu().numpy(), c)
        before = [dc.clone() for dc in t.device_cores]
        nrm = t.norm()
        assert abs(nrm - tt_ops.norm(cl)) <= 1e-13 * tt_ops.norm(cl)
        for b, dc in zip(before, t.device_cores):
            assert torch.equal(b, dc)


def test_evp_adaptive_tolerance_matches_fixed(env):
    """Inexact ALS (adaptive local tolerance) stays within 1e-10 of the
    fixed-tolerance eigenvalue and needs fewer Krylov steps."""
    from scikit_tt_amd.tensor_train import from_host_batch
    import scikit_tt_amd.solvers.evp as evp
    lists = []
    for s in range(4):
        H, y0, J, g = W.tfim_chain(seed=30 + s, d=16, rank=8)
        lists += [H, W.right_orthonormal(y0)]
    tts = from_host_batch(lists)
    ops, xs = tts[0::2], tts[1::2]
    kw = dict(repeats=6, conv_eps=0.0, solver='eig', sigma=-100.0, return_info=True)
    fixed, fi = evp.als_batch(ops, xs, **kw)
    adapt, ai = evp.als_batch(ops, xs, adaptive_tol=True, **kw)
    for f, a_, i1, i2 in zip(fixed, adapt, fi, ai):
        assert i1.status == 0 and i2.status == 0
        assert abs(f[0] - a_[0]) <= 1e-10 * abs(f[0])
    assert sum(i.krylov_iters for i in ai) < sum(i.krylov_iters for i in fi)
    # the policy applies to that call only
    again, _ = evp.als_batch(ops[:1], xs[:1], **kw)
    assert again[0][0] == fixed[0][0]


def test_evp_batch_rerun_bit_identical(env):
    """The engine is deterministic: the same batch twice gives bit-identical
    eigenvalues, ranks and cores (a data race between the waves of a system
    shows up here as run-to-run differences)."""
    from scikit_tt_amd.tensor_train import from_host_batch
    import scikit_tt_amd.solvers.evp as evp
    lists = []
    for s in range(128):
        H, y0, J, g = W.tfim_chain(seed=100 + s, d=16, rank=12)
        lists += [H, W.right_orthonormal(y0)]
    tts = from_host_batch(lists)
    ops, xs = tts[0::2], tts[1::2]
    kw = dict(repeats=4, conv_eps=0.0, solver='eig', sigma=-100.0)
    first = evp.als_batch(ops, xs, **kw)
    for rep in range(2):
        again = evp.als_batch(ops, xs, **kw)
        for (l1, t1, _), (l2, t2, _) in zip(first, again):
            assert l1 == l2
            assert t1.ranks == t2.ranks
            for c1, c2 in zip(t1.device_cores, t2.device_cores):
                assert torch.equal(c1, c2)


def test_sle_laplace_large_full_size(env):
    """C3 at its real size (d=20, n=32, TT rank 32, N = 32768 per interior
    local problem, global-memory path): two sweeps reach rel.res 2e-9
    (the rank-32 approximation levels off near 1e-9)."""
    from scikit_tt_amd.tensor_train import from_host_batch, residual_error
    import scikit_tt_amd.solvers.sle as sle
    A, x0, b = W.laplace_large()
    tA, tx, tb = from_host_batch([A, x0, b])
    (x,), (info,) = sle.als_batch([tA], [tx], [tb], repeats=2, return_info=True)
    assert info.status == 0 and info.unconverged == 0
    assert x.ranks == [1] + [32] * 19 + [1]
    rr = residual_error(tA, x, tb) / tb.norm()
    assert rr <= 2e-9, rr


def test_sle_uneven_batch_full_size(env):
    """C4 at its real size: 512 uneven shifted-identity systems in one launch.
    Every system succeeds, the residual falls, and sample systems give the
    same cores alone as in the batch."""
    from scikit_tt_amd.tensor_train import from_host_batch, residual_error
    import scikit_tt_amd.solvers.sle as sle
    insts = [W.shifted_batch_instance(s) for s in range(512)]
    tts = from_host_batch([c for inst in insts for c in inst])
    As, xs, bs = tts[0::3], tts[1::3], tts[2::3]
    sols, infos = sle.als_batch(As, xs, bs, repeats=2, return_info=True)
    assert all(i.status == 0 for i in infos)
    for s in (0, 17, 255, 511):
        r0 = residual_error(As[s], xs[s], bs[s])
        r1 = residual_error(As[s], sols[s], bs[s])
        assert r1 < 1e-2 * r0, (s, r0, r1)
        (alone,), _ = sle.als_batch([As[s]], [xs[s]], [bs[s]], repeats=2, return_info=True)
        for c1, c2 in zip(alone.device_cores, sols[s].device_cores):
            assert torch.equal(c1, c2)


def _nonsym_slices_symmetric_operator(rng, d, n, scale=0.3):
    """A = sum over k of B_k (x) B_k^T-chains: a rank-2 TT operator whose
    slices are not symmetric one by one but whose whole matrix is symmetric
    (core i = [[C_i, 0], [0, C_i^T]] with first/last cores closing the two
    paths), plus a diagonal shift so the spectrum has a clear bottom."""
    cores = []
    for i in range(d):
        C = scale * rng.standard_normal((n, n))
        if i == 0:
            c = np.zeros((1, n, n, 3))
            c[0, :, :, 0], c[0, :, :, 1], c[0, :, :, 2] = C, C.T, np.eye(n)
        elif i == d - 1:
            c = np.zeros((3, n, n, 1))
            c[0, :, :, 0], c[1, :, :, 0], c[2, :, :, 0] = C, C.T, np.diag(np.arange(n, dtype=float))
        else:
            c = np.zeros((3, n, n, 3))
            c[0, :, :, 0], c[1, :, :, 1], c[2, :, :, 2] = C, C.T, np.eye(n)
        cores.append(c)
    return cores


def test_evp_nonsymmetric_operator_raises(env):
    """evp.als with an operator that is not symmetric never returns the
    eigenpair of a symmetric solver: 'eigh' raises NotImplementedError; 'eig'
    runs the general local eigensolver, whose selected eigenvalues here are
    complex (the operator's nearest to sigma are), which raises too (the
    reference would keep their real parts and complex eigenvectors,
    scikit_tt/solvers/evp.py:424-443)."""
    import scikit_tt_amd.solvers.evp as evp
    rng = np.random.default_rng(11)
    A = W.random_tt(rng, [3] * 4, [1, 2, 2, 2, 1], col_dims=[3] * 4)
    x0 = W.right_orthonormal(W.random_tt(rng, [3] * 4, 2))
    with pytest.raises(NotImplementedError):
        evp.als(_tt(A), _tt(x0), repeats=2, conv_eps=0, solver='eig', sigma=-100)
    with pytest.raises(NotImplementedError):
        evp.als(_tt(A), _tt(x0), repeats=2, conv_eps=0, solver='eigh')


def test_evp_symmetric_operator_with_nonsymmetric_slices(env):
    """The symmetry test looks at the whole operator: slices that are not
    symmetric one by one but add up to a symmetric matrix are accepted, and
    the eigenvalue matches the oracle's (dense local eigensolves)."""
    import scikit_tt_amd.solvers.evp as evp
    rng = np.random.default_rng(12)
    A = _nonsym_slices_symmetric_operator(rng, 4, 3)
    M = tt_ops.matricize(A)
    assert np.abs(M - M.T).max() < 1e-12 * np.abs(M).max()
    assert any(np.abs(c - c.transpose(0, 2, 1, 3)).max() > 0.1 for c in A)
    x0 = W.right_orthonormal(W.random_tt(rng, [3] * 4, 3))
    lam, y, it = evp.als(_tt(A), _tt(x0), repeats=3, conv_eps=0, solver='eig', sigma=-100)
    lr, yr, itr = evp_als.als(A, x0, repeats=3, conv_eps=0, solver='eig', sigma=-100)
    assert it == itr
    assert abs(lam - lr) < 1e-10 * max(1.0, abs(lr)), (lam, lr)


def test_output_descriptor_capacity_checked(env):
    """The C ABI rejects output buffers smaller than the initial guess (no
    out-of-bounds device writes for direct C callers)."""
    lib, ctx, _lib = env
    from scikit_tt_amd.tensor_train import TT
    A, x0, b = W.laplace_small(d=4, n=3, rank=3)
    tA, tx, tb = _tt(A), _tt(x0), _tt(b)
    small = W.random_tt(np.random.default_rng(1), [3] * 4, 2)
    ts = _tt(small)
    res = (_lib.Result * 1)()
    packs = [t._desc() for t in (tA, tx, tb, ts)]  # the descriptors' arrays stay alive through the call
    rc = lib.ttals_sle_als_batch(ctx, None, 1, *[ctypes.byref(p.desc) for p in packs], 1, 0, 1e-12, 100, res)
    assert rc == _lib.E_SHAPE
    assert b'capacity' in lib.ttals_last_error(ctx)
