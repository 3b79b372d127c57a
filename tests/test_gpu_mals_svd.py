"""GPU: sle.mals, the TT-SVD constructor, tt2qtt and truncated_svd against
vectors the reference itself wrote (tests/golden/sle_mals_toeplitz_d10.npz,
tests/golden/tt_svd_qtt.npz) and against LAPACK.

* truncated_svd: singular values to 1e-12 relative of LAPACK's, orthonormal
  factors, reconstruction, the kept rank by the reference's rule;
* TT(full, threshold, max_rank): ranks identical, the tensor to 1e-10
  relative (untruncated) / the reference's truncated tensor to 1e-8;
* tt2qtt: ranks identical, tensor to 1e-10;
* mals on the reference's Toeplitz test system (tests/test_sle.py:61-85):
  ranks identical, the iterate to 1e-7 of the reference's, and the
  reference test's own bar against the dense solution (1e-7).
"""
import numpy as np
import pytest

from oracle import tt_ops
from tests.conftest import load_golden

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def env():
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    from scikit_tt_amd import _lib
    torch.cuda.set_device(0)
    _lib.load()


def _tt(cores):
    from scikit_tt_amd.tensor_train import TT
    return TT([np.asarray(c) for c in cores])


def _reldiff(x, y):
    return tt_ops.norm(tt_ops.sub(x, y)) / tt_ops.norm(y)


@pytest.mark.parametrize('shape', [(40, 12), (12, 40), (30, 30), (200, 7)])
@pytest.mark.parametrize('thr,maxr', [(0, np.inf), (1e-3, np.inf), (0, 5), (1e-6, 4)])
def test_truncated_svd(env, shape, thr, maxr):
    from scikit_tt_amd.utils import truncated_svd
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    m, n = shape
    k0 = min(m, n)
    # gapped spectrum so the kept rank is well defined
    Uq, _ = np.linalg.qr(rng.standard_normal((m, k0)))
    Vq, _ = np.linalg.qr(rng.standard_normal((n, k0)))
    sv = 10.0 ** -np.arange(k0, dtype=float) * 0.9 ** np.arange(k0)
    Y = (Uq * sv) @ Vq.T
    u, s, v = truncated_svd(Y, threshold=thr, max_rank=maxr)
    u, s, v = u.cpu().numpy(), s.cpu().numpy(), v.cpu().numpy()
    ref = np.linalg.svd(Y, compute_uv=False)
    keep = np.nonzero(ref / ref[0] > thr)[0] if thr != 0 else np.arange(len(ref))
    k = len(keep) if maxr == np.inf else min(len(keep), int(maxr))
    assert len(s) == k
    np.testing.assert_allclose(s, ref[:k], rtol=1e-10, atol=1e-14 * ref[0])
    np.testing.assert_allclose(u.T @ u, np.eye(k), atol=1e-12)
    big = s > 1e-10 * s[0]
    np.testing.assert_allclose((v @ v.T)[np.ix_(big, big)], np.eye(int(big.sum())), atol=1e-10)
    full = (u * s) @ v
    best = (np.linalg.svd(Y)[0][:, :k] * ref[:k]) @ np.linalg.svd(Y)[2][:k]
    assert np.linalg.norm(full - best) <= 1e-10 * np.linalg.norm(Y)


def test_tt_svd_constructor_vs_reference(env):
    from scikit_tt_amd.tensor_train import TT
    g = load_golden('tt_svd_qtt')
    full = g['full']
    for tag, kw, shape in (('exact', {}, [2] * 12), ('thr', {'threshold': 1e-10}, [2] * 12),
                           ('maxr', {'max_rank': 5}, [4] * 6)):
        t = TT(full.reshape(shape), **kw)
        assert t.ranks == list(g['svd_ranks_' + tag]), tag
        ref = g['tt']('svd_' + tag)
        assert _reldiff(t.cores[:], ref) < (1e-8 if tag == 'maxr' else 1e-10), tag
    with pytest.raises(ValueError):
        TT(np.zeros((2, 2, 2)))


def test_tt2qtt_vs_reference(env):
    g = load_golden('tt_svd_qtt')
    base = _tt(g['tt']('q_in'))
    for tag, thr in (('q_out', 0), ('q_out_thr', 1e-2)):
        q = base.tt2qtt([[2, 2], [2, 2, 2], [4]], [[2, 2], [2, 1, 1], [2, 2]], threshold=thr)
        assert q.ranks == list(g['q_ranks' if tag == 'q_out' else 'q_ranks_thr'])
        assert _reldiff(q.cores[:], g['tt'](tag)) < 1e-10


@pytest.mark.parametrize('solver', ['solve', 'lu'])
def test_mals_vs_reference(env, solver):
    import scikit_tt_amd.solvers.sle as sle
    g = load_golden('sle_mals_toeplitz_d10')
    A, x0, b = _tt(g['tt']('A')), _tt(g['tt']('x0')), _tt(g['tt']('b'))
    x = sle.mals(A, x0, b, repeats=1, solver=solver, threshold=1e-14, max_rank=10)
    assert x.ranks == list(g['ranks_' + solver])
    assert _reldiff(x.cores[:], g['tt']('x_' + solver)) < 1e-7
    ds = g['dense_solution']
    assert np.linalg.norm(x.matricize() - ds) / np.linalg.norm(ds) < 1e-7
