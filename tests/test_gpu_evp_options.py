"""GPU: evp.als with deflation (previous/shift), a generalised problem
(operator_gevp), several eigenpairs (number_ev > 1) and non-symmetric
operators, against vectors the reference itself wrote
(tests/golden/evp_options.npz, make_golden.py 'evp_options'): the
reference's own evp test problem (the Ulam operator of tests/test_evp.py: as
it is with 'eig', one pair with operator_gevp = I and three pairs; its
symmetric part with 3 pairs 'eigh') and a TFIM chain.  Eigenvalues to 1e-9
relative, iteration counts and ranks exactly, eigentensors up to sign to 1e-6
(the local eigenvectors and SVD factors are fixed up to sign only)."""
import numpy as np
import pytest

from oracle import tt_ops
from tests.conftest import load_golden

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def g():
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    from scikit_tt_amd import _lib
    torch.cuda.set_device(0)
    _lib.load()
    return load_golden('evp_options')


def _tt(cores):
    from scikit_tt_amd.tensor_train import TT
    return TT([np.asarray(c) for c in cores])


def _close_up_to_sign(y, ref, tol):
    d1 = tt_ops.norm(tt_ops.sub(y.cores[:], ref))
    d2 = tt_ops.norm(tt_ops.add(y.cores[:], ref))
    assert min(d1, d2) <= tol * tt_ops.norm(ref), (d1, d2)


def test_several_eigenpairs_reference_test_problem(g):
    import scikit_tt_amd.solvers.evp as evp
    print('case ulam number_ev=3 eigh', flush=True)
    vals, vecs, it = evp.als(_tt(g['tt']('ulam_A')), _tt(g['tt']('ulam_x0')), number_ev=3, repeats=10,
                             solver='eigh')
    ref = g['ulam_vals']
    print('  eigenvalues', vals, 'reference', ref, flush=True)
    np.testing.assert_allclose(vals, ref, rtol=1e-9)
    assert it == int(g['ulam_iters'])
    for e in range(3):
        assert vecs[e].ranks == list(g['ulam_ranks'])
        _close_up_to_sign(vecs[e], g['tt']('ulam_vec%d' % e), 1e-6)
    # the reference test's own tolerance against the dense eigenvalues
    assert np.all(np.abs(vals - g['ulam_dense_top3']) <= 5e-3 * np.abs(g['ulam_dense_top3']))


def test_generalised_problem(g):
    import scikit_tt_amd.solvers.evp as evp
    print('case tfim operator_gevp', flush=True)
    lam, y, it = evp.als(_tt(g['tt']('tfim_A')), _tt(g['tt']('tfim_x0')), operator_gevp=_tt(g['tt']('tfim_M')),
                         repeats=6, solver='eig', sigma=-100.0)
    ref = float(g['gevp_val'])
    print('  eigenvalue %.15f reference %.15f dense %.15f' % (lam, ref, float(g['gevp_dense'])), flush=True)
    assert abs(lam - ref) <= 1e-9 * abs(ref)
    assert it == int(g['gevp_iters'])
    _close_up_to_sign(y, g['tt']('gevp_vec'), 1e-6)


def test_deflation_of_the_ground_state(g):
    import scikit_tt_amd.solvers.evp as evp
    print('case tfim previous/shift', flush=True)
    lam, y, it = evp.als(_tt(g['tt']('tfim_A')), _tt(g['tt']('tfim_x1')), previous=[_tt(g['tt']('defl_prev'))],
                         shift=50.0, repeats=6, solver='eig', sigma=-100.0)
    ref = float(g['defl_val'])
    print('  eigenvalue %.15f reference %.15f dense %s' % (lam, ref, g['dense_lowest2']), flush=True)
    assert abs(lam - ref) <= 1e-9 * abs(ref)
    assert it == int(g['defl_iters'])
    _close_up_to_sign(y, g['tt']('defl_vec'), 1e-6)


def test_two_eigenpairs_eig(g):
    import scikit_tt_amd.solvers.evp as evp
    print('case tfim number_ev=2 eig', flush=True)
    vals, vecs, it = evp.als(_tt(g['tt']('tfim_A')), _tt(g['tt']('tfim_x0')), number_ev=2, repeats=6, solver='eig',
                             sigma=-100.0)
    print('  eigenvalues', vals, 'reference', g['two_vals'], flush=True)
    np.testing.assert_allclose(vals, g['two_vals'], rtol=1e-9)
    assert it == int(g['two_iters'])
    for e in range(2):
        _close_up_to_sign(vecs[e], g['tt']('two_vec%d' % e), 1e-6)


def test_non_spd_gevp_and_nonsymmetric_raise(g):
    import scikit_tt_amd.solvers.evp as evp
    A = _tt(g['tt']('tfim_A'))
    x0 = _tt(g['tt']('tfim_x0'))
    negM = [-np.asarray(c) if i == 0 else np.asarray(c) for i, c in enumerate(g['tt']('tfim_M'))]
    with pytest.raises(NotImplementedError):
        evp.als(A, x0, operator_gevp=_tt(negM), repeats=1)
    rng = np.random.default_rng(1)
    skew = [np.asarray(c) + (0.1 * rng.standard_normal(np.asarray(c).shape) if i == 2 else 0.0)
            for i, c in enumerate(g['tt']('tfim_A'))]
    with pytest.raises(NotImplementedError):  # 'eigh' of a non-symmetric operator
        evp.als(_tt(skew), x0, number_ev=2, repeats=1, solver='eigh')


def test_nonsymmetric_reference_test_problem_one_pair(g):
    """tests/test_evp.py:test_als_eig, first call: the Ulam operator as it is
    (not symmetric), operator_gevp = identity, 'eig' nearest sigma = 1: the
    general local eigensolver (Hessenberg QR, inverse iteration)."""
    import scikit_tt_amd.solvers.evp as evp
    from scikit_tt_amd import tensor_train as tt
    print('case ulam (not symmetric) eig, operator_gevp = I', flush=True)
    A = _tt(g['tt']('ulam_ns_A'))
    lam, y, it = evp.als(A, _tt(g['tt']('ulam_x0')), operator_gevp=tt.eye(A.row_dims), repeats=10)
    ref = float(g['ulam_ns_val'])
    print('  eigenvalue %.15f reference %.15f' % (lam, ref), flush=True)
    assert abs(lam - ref) <= 1e-9 * abs(ref)
    assert it == int(g['ulam_ns_iters'])
    _close_up_to_sign(y, g['tt']('ulam_ns_vec'), 1e-6)


def test_nonsymmetric_reference_test_problem_three_pairs(g):
    """tests/test_evp.py:test_als_eig, second call: number_ev = 3, 'eig'."""
    import scikit_tt_amd.solvers.evp as evp
    print('case ulam (not symmetric) eig, number_ev = 3', flush=True)
    vals, vecs, it = evp.als(_tt(g['tt']('ulam_ns_A')), _tt(g['tt']('ulam_x0')), number_ev=3, repeats=10)
    print('  eigenvalues', vals, 'reference', g['ulam_ns3_vals'], flush=True)
    np.testing.assert_allclose(vals, g['ulam_ns3_vals'], rtol=1e-9)
    assert it == int(g['ulam_ns3_iters'])
    for e in range(3):
        _close_up_to_sign(vecs[e], g['tt']('ulam_ns3_vec%d' % e), 1e-6)
    # the reference test's tolerance against the dense eigenvalues
    assert np.all(np.abs(vals - g['ulam_ns_dense_top3']) <= 5e-3 * np.abs(g['ulam_ns_dense_top3']))
