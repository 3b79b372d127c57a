"""Seeded synthetic inputs for the TT-ALS workloads (host numpy cores).

All builders return plain lists of float64 numpy cores ``(r, m, n, r')`` so
that the same arrays can be handed to the device ``TT`` class and to a CPU
checker.  Random cores are drawn with ``numpy.random.default_rng(seed)`` in
core order i = 0..d-1 with ``standard_normal(shape)``.

Workloads (sizes as used by the tests and the benchmark):

* ``laplace``: -Laplacian on a uniform grid, d modes of size n, rank-2
  Kronecker-sum operator (layout of examples/tensorkrylov.py:127-140), all-ones
  or random right-hand side.
* ``tfim``: inhomogeneous transverse-field Ising chain, MPO of rank 3 in the
  structure of scikit_tt/models.py:31-40 (S_i = -g_i X, coupling -J_i Z Z).
* ``shifted``: alpha * I + rank-3 random perturbation (TT sum, rank 4).
"""
import numpy as np

__all__ = [
    'laplace_1d', 'laplace_operator', 'tfim_operator', 'tfim_ground_energy',
    'identity_operator', 'random_tt', 'ones_tt', 'shifted_operator', 'tt_sum',
    'laplace_small', 'tfim_chain', 'laplace_large', 'shifted_batch_instance', 'right_orthonormal',
    'laplace_lowest_eigenvalue', 'serving_request',
]


def laplace_1d(n):
    """(n+1)^2 * tridiag(-1, 2, -1): the 1-D Dirichlet Laplacian, h = 1/(n+1)."""
    h2 = float(n + 1) ** 2
    return h2 * (2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1))


def laplace_operator(d, n):
    """Rank-2 TT operator of -Laplacian = sum_k I x .. x D_k x .. x I.

    core 0: [[D, I]], core k: [[I, 0], [D, I]], core d-1: [[I], [D]]
    (the Kronecker-sum layout of examples/tensorkrylov.py:127-140).
    """
    D, I = laplace_1d(n), np.eye(n)
    first = np.zeros((1, n, n, 2))
    first[0, :, :, 0], first[0, :, :, 1] = D, I
    mid = np.zeros((2, n, n, 2))
    mid[0, :, :, 0], mid[1, :, :, 0], mid[1, :, :, 1] = I, D, I
    last = np.zeros((2, n, n, 1))
    last[0, :, :, 0], last[1, :, :, 0] = I, D
    if d == 1:
        return [D.reshape(1, n, n, 1)]
    return [first] + [mid.copy() for _ in range(d - 2)] + [last]


def identity_operator(dims, alpha=1.0):
    """alpha * I as a rank-1 TT operator (alpha folded into core 0)."""
    cores = [np.eye(n).reshape(1, n, n, 1) for n in dims]
    cores[0] = alpha * cores[0]
    return cores


def tt_sum(x, y):
    """TT sum with block-diagonal rank concatenation (x block first).

    Same rank layout as scikit_tt/tensor_train.py:313-333.
    """
    d = len(x)
    rx = [c.shape[0] for c in x] + [1]
    ry = [c.shape[0] for c in y] + [1]
    rs = [1] + [rx[i] + ry[i] for i in range(1, d)] + [1]
    out = []
    for i in range(d):
        c = np.zeros((rs[i], x[i].shape[1], x[i].shape[2], rs[i + 1]))
        c[:x[i].shape[0], :, :, :x[i].shape[3]] = x[i]
        c[rs[i] - y[i].shape[0]:, :, :, rs[i + 1] - y[i].shape[3]:] = y[i]
        out.append(c)
    return out


def random_tt(rng, row_dims, ranks, col_dims=None, scale=1.0):
    """Cores with i.i.d. N(0, scale^2) entries, drawn in core order."""
    d = len(row_dims)
    if not isinstance(ranks, (list, tuple)):
        ranks = [1] + [int(ranks)] * (d - 1) + [1]
    col_dims = [1] * d if col_dims is None else col_dims
    return [scale * rng.standard_normal((ranks[i], row_dims[i], col_dims[i], ranks[i + 1]))
            for i in range(d)]


def ones_tt(dims, rank=1):
    """All-ones TT of the given rank (scikit_tt/tensor_train.py:1838-1867)."""
    d = len(dims)
    r = [1] + [rank] * (d - 1) + [1]
    return [np.ones((r[i], dims[i], 1, r[i + 1])) for i in range(d)]


def tfim_operator(J, g):
    """H = -sum_i J_i Z_i Z_{i+1} - sum_i g_i X_i as a rank-3 MPO.

    Same automaton as scikit_tt/models.py:31-40 with site-dependent
    couplings: core 0 [[S_0, Z, I]], core i [[I,0,0],[M_i,0,0],[S_i,Z,I]],
    last core [I, M_{d-1}, S_{d-1}]^T with S_i = -g_i X and M_i = -J_{i-1} Z.
    """
    d = len(g)
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    Z = np.array([[1.0, 0.0], [0.0, -1.0]])
    I = np.eye(2)
    cores = []
    for i in range(d):
        S = -g[i] * X
        M = -J[i - 1] * Z if i > 0 else None
        if i == 0:
            c = np.zeros((1, 2, 2, 3))
            c[0, :, :, 0], c[0, :, :, 1], c[0, :, :, 2] = S, Z, I
        elif i == d - 1:
            c = np.zeros((3, 2, 2, 1))
            c[0, :, :, 0], c[1, :, :, 0], c[2, :, :, 0] = I, M, S
        else:
            c = np.zeros((3, 2, 2, 3))
            c[0, :, :, 0] = I
            c[1, :, :, 0] = M
            c[2, :, :, 0], c[2, :, :, 1], c[2, :, :, 2] = S, Z, I
        cores.append(c)
    return cores


def tfim_ground_energy(J, g):
    """Exact ground energy of the open TFIM chain (free fermions).

    E0 = -sum of the singular values of the bidiagonal matrix with diagonal g
    and super-diagonal J (Jordan-Wigner + Bogoliubov; Lieb-Schultz-Mattis).
    """
    d = len(g)
    B = np.diag(np.asarray(g, dtype=float))
    if d > 1:
        B += np.diag(np.asarray(J, dtype=float), k=1)
    return -float(np.sum(np.linalg.svd(B, compute_uv=False)))


def shifted_operator(rng, d, n, alpha=None, rank=3, sigma=0.1):
    """alpha*I + P with P a rank-`rank` TT operator of N(0, sigma^2) cores."""
    alpha = float(d) if alpha is None else alpha
    P = random_tt(rng, [n] * d, rank, col_dims=[n] * d, scale=sigma)
    return tt_sum(identity_operator([n] * d, alpha), P)


# ---------------------------------------------------------------------------
# named workloads


def laplace_small(seed=0, d=10, n=8, rank=5):
    """-Laplacian d=10, n=8 (rank 2); b = ones (rank 1); x0 N(0,1) rank 5 (rng(seed))."""
    A = laplace_operator(d, n)
    b = ones_tt([n] * d, 1)
    x0 = random_tt(np.random.default_rng(seed), [n] * d, rank)
    return A, x0, b


def tfim_chain(seed=0, d=32, rank=16):
    """TFIM chain d=32, J,g ~ U[0.5,1.5] (rng(seed)), raw N(0,1) x0 of rank 16.

    Draw order: J (d-1 values), g (d values), then the x0 cores.  The caller
    right-orthonormalises and normalises x0 before evp.als, as the reference
    tests do (tests/test_ode.py:143-145, tests/test_evp.py:45).
    """
    rng = np.random.default_rng(seed)
    J = rng.uniform(0.5, 1.5, d - 1)
    g = rng.uniform(0.5, 1.5, d)
    x0 = random_tt(rng, [2] * d, rank)
    return tfim_operator(J, g), x0, J, g


def laplace_large(d=20, n=32, rank=32, rhs_rank=4, rhs_seed=1, x_seed=2):
    """-Laplacian d=20, n=32; b N(0,1) rank 4 (rng(1)); x0 N(0,1) rank 32 (rng(2))."""
    A = laplace_operator(d, n)
    b = random_tt(np.random.default_rng(rhs_seed), [n] * d, rhs_rank)
    x0 = random_tt(np.random.default_rng(x_seed), [n] * d, rank)
    return A, x0, b


def shifted_batch_instance(s, d_range=(8, 24), n_choices=(4, 8, 16), r_range=(4, 48),
                           rhs_rank=4):
    """Instance s of the uneven batch: rng(s) draws d, n, r, then P, b, x0.

    d ~ U{8..24}, n in {4, 8, 16}, r ~ U{4..48}; A = d*I + rank-3 N(0, 0.1^2)
    perturbation (TT rank 4), b rank 4 N(0,1), x0 rank r N(0,1).
    """
    rng = np.random.default_rng(s)
    d = int(rng.integers(d_range[0], d_range[1] + 1))
    n = int(rng.choice(n_choices))
    r = int(rng.integers(r_range[0], r_range[1] + 1))
    A = shifted_operator(rng, d, n)
    b = random_tt(rng, [n] * d, rhs_rank)
    x0 = random_tt(rng, [n] * d, r)
    return A, x0, b


def right_orthonormal(cores, normalize=True):
    """The same tensor with cores 1..d-1 right-orthonormal (host QR sweep).

    From the last core down: the transposed unfolding (n r' x r) of core i is
    factored Q R, core i becomes Q^T (rank min(r, n r')) and R^T is absorbed
    into core i-1.  With ``normalize`` the first core is scaled to unit
    Frobenius norm, which is then the norm of the whole tensor.
    """
    out = [np.array(c, dtype=np.float64) for c in cores]
    for i in range(len(out) - 1, 0, -1):
        r, m, n, r2 = out[i].shape
        q, rf = np.linalg.qr(out[i].reshape(r, m * n * r2).T)
        k = q.shape[1]
        out[i] = q.T.reshape(k, m, n, r2)
        out[i - 1] = np.tensordot(out[i - 1], rf.T, axes=([3], [0]))
    if normalize:
        out[0] = out[0] / np.linalg.norm(out[0])
    return out



def laplace_lowest_eigenvalue(d, n):
    """Lowest eigenvalue of the d-mode -Laplacian of laplace_operator: d times
    the lowest eigenvalue 4 (n+1)^2 sin^2(pi / (2 (n+1))) of the 1-D factor
    (a Kronecker sum's eigenvalues are sums of the factors')."""
    return d * 4.0 * (n + 1) ** 2 * np.sin(np.pi / (2 * (n + 1))) ** 2


def serving_request(k, d_range=(8, 64), n_choices=(2, 4, 8, 16), r_range=(8, 128)):
    """Request k of the mixed serving workload (rng(k)): sle or evp with equal
    probability, d ~ U{8..64}, n in {2, 4, 8, 16}, rank log-uniform in [8, 128].

    sle: the -Laplacian with a rank-4 N(0,1) rhs (the large-Laplacian
    configuration's form) or d*I + a rank-3 N(0, 0.1^2) perturbation with a
    rank-4 rhs (the uneven-batch form), equally likely; x0 rank r N(0,1).
    evp: the TFIM chain for n = 2, else the -Laplacian; x0 rank r N(0,1),
    right-orthonormal, unit norm; the lowest eigenvalue ('eig', sigma=-100).
    Returns a dict with 'kind', the host cores and the sizes.
    """
    rng = np.random.default_rng(k)
    kind = 'sle' if rng.random() < 0.5 else 'evp'
    d = int(rng.integers(d_range[0], d_range[1] + 1))
    n = int(rng.choice(n_choices))
    r = int(round(float(np.exp(rng.uniform(np.log(r_range[0]), np.log(r_range[1]))))))
    out = {'kind': kind, 'd': d, 'n': n, 'rank': r}
    if kind == 'sle':
        if rng.random() < 0.5:
            out['family'] = 'laplace'
            out['A'] = laplace_operator(d, n)
        else:
            out['family'] = 'shifted'
            out['A'] = shifted_operator(rng, d, n)
        out['b'] = random_tt(rng, [n] * d, 4)
        out['x0'] = random_tt(rng, [n] * d, r)
    else:
        if n == 2:
            out['family'] = 'tfim'
            J = rng.uniform(0.5, 1.5, d - 1)
            g = rng.uniform(0.5, 1.5, d)
            out['A'] = tfim_operator(J, g)
            out['exact'] = tfim_ground_energy(J, g)
        else:
            out['family'] = 'laplace'
            out['A'] = laplace_operator(d, n)
            out['exact'] = laplace_lowest_eigenvalue(d, n)
        out['x0'] = right_orthonormal(random_tt(rng, [n] * d, r))
    return out
