"""GPU: the engine's algorithmic byte count (``SolveInfo.bytes``).

The count follows the per-Krylov-iteration formula of the local matvec: the
left and right stacks, the operator core and the iterate read once
(8 (aPa + anb + PnnQ + bQb) bytes), plus the input and output vectors (16 N)
where they live in global memory; the solvers' vector work on global-memory
vectors adds one element moved per vector flop (so at most 8 bytes per
algorithmic flop of the solve).  A dense local problem (N <= 96) is charged
the operator pieces read once and its micro matrix written and read once:
8 (aPa + PnnQ + bQb) + 16 N^2.

* all-dense system: the count equals the sum of the dense charge over the
  positions of each sweep (forward 0..d-2, backward d-1..0), recomputed here
  from the shapes;
* dense ends with Krylov interior positions: the dense part exactly, the
  Krylov part at least the Krylov iterations times the smallest per-matvec
  charge and at most (iterations + 3 per local solve) times the largest plus
  8 bytes per flop of the solve;
* a gang (every local problem in global memory): iterations + 1 or 2
  matvecs per local solve (the initial residual, and the confirming residual
  unless GMRES converged in one or two steps) times the per-matvec charge
  with the 16 N vector term, between the smallest and largest position, plus
  at most 8 bytes per flop of vector work.
"""
import numpy as np
import pytest

from scikit_tt_amd import workloads as W

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def env():
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    from scikit_tt_amd import _lib
    torch.cuda.set_device(0)
    _lib.load()


def _shapes(A, x):
    """(a, P, n, Q, b) of every position."""
    return [(x[i].shape[0], A[i].shape[0], x[i].shape[1], A[i].shape[3], x[i].shape[3]) for i in range(len(x))]


def _mv(a, P, n, Q, b, global_vec):
    N = a * n * b
    return 8.0 * (a * P * a + N + P * n * n * Q + b * Q * b + (2 * N if global_vec else 0))


def _dense(a, P, n, Q, b):
    N = a * n * b
    return 8.0 * (a * P * a + P * n * n * Q + b * Q * b) + 16.0 * N * N


def _positions(d, repeats):
    return (list(range(d - 1)) + list(range(d - 1, -1, -1))) * repeats


def _solve(systems, repeats, gang=False):
    import scikit_tt_amd.solvers.sle as sle
    from scikit_tt_amd.tensor_train import TT
    As = [TT([np.asarray(c) for c in A]) for A, _, _ in systems]
    Xs = [TT([np.asarray(c) for c in x]) for _, x, _ in systems]
    Bs = [TT([np.asarray(c) for c in b]) for _, _, b in systems]
    if gang:
        _, infos = sle.als_gang(As, Xs, Bs, repeats=repeats, gang=4, return_info=True)
    else:
        _, infos = sle.als_batch(As, Xs, Bs, repeats=repeats, return_info=True)
    arr = infos.array
    assert (arr['status'] == 0).all()
    return arr


def test_bytes_dense_positions_exact(env):
    print('case bytes: all-dense systems d=5 n=4 r=3, 2 sweeps', flush=True)
    systems = [W.shifted_batch_instance(s, d_range=(5, 5), n_choices=(4,), r_range=(3, 3)) for s in range(3)]
    arr = _solve(systems, 2)
    for s, (A, x, _) in enumerate(systems):
        sh = _shapes(A, x)
        assert all(a * n * b <= 96 for a, _, n, _, b in sh)
        want = sum(_dense(*sh[i]) for i in _positions(len(x), 2))
        assert arr['local_solves'][s] == len(_positions(len(x), 2))
        assert arr['bytes'][s] == pytest.approx(want, rel=1e-12), (s, arr['bytes'][s], want)


def test_bytes_krylov_positions_bounded(env):
    print('case bytes: d=4 n=8 r=8 (dense ends, Krylov interior), 1 sweep', flush=True)
    systems = [W.shifted_batch_instance(s, d_range=(4, 4), n_choices=(8,), r_range=(8, 8)) for s in range(2)]
    arr = _solve(systems, 1)
    for s, (A, x, _) in enumerate(systems):
        sh = _shapes(A, x)
        pos = _positions(len(x), 1)
        dense = [i for i in pos if sh[i][0] * sh[i][2] * sh[i][4] <= 96]
        kry = [i for i in pos if sh[i][0] * sh[i][2] * sh[i][4] > 96]
        assert dense and kry
        kit = int(arr['krylov_iters'][s])
        part = arr['bytes'][s] - sum(_dense(*sh[i]) for i in dense)
        lo = kit * min(_mv(*sh[i], False) for i in kry)
        hi = (kit + 3 * len(kry)) * max(_mv(*sh[i], True) for i in kry) + 8.0 * arr['flops'][s]
        assert lo <= part <= hi, (s, lo, part, hi)


def test_bytes_gang_bounded(env):
    print('case bytes: one system d=6 n=8 r=8 on a gang of 4, 1 sweep', flush=True)
    systems = [W.shifted_batch_instance(7, d_range=(6, 6), n_choices=(8,), r_range=(8, 8))]
    arr = _solve(systems, 1, gang=True)
    A, x, _ = systems[0]
    sh = _shapes(A, x)
    pos = _positions(len(x), 1)
    kit = int(arr['krylov_iters'][0])
    assert arr['local_solves'][0] == len(pos)
    lo = (kit + len(pos)) * min(_mv(*sh[i], True) for i in pos)
    hi = (kit + 2 * len(pos)) * max(_mv(*sh[i], True) for i in pos) + 8.0 * arr['flops'][0]
    assert lo <= arr['bytes'][0] <= hi, (lo, arr['bytes'][0], hi)
