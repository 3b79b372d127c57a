"""GPU: callers of sle.als against vectors the reference itself wrote.

* ode.implicit_euler / trapezoidal_rule / adaptive_step_size and the error
  functions on the reference's ODE test operator (signaling cascade,
  quantised, tests/test_ode.py:176-192), fewer steps
  (tests/golden/ode_signaling_cascade.npz): final trains to 1e-7 relative,
  step errors to 1e-3 relative above 1e-12 (they are small residual norms:
  the reference computes them by an SVD-based TT norm that is itself only
  good to ~1e-4 relative at residuals ~1e-8, the device by a QR sweep), the
  adaptive method's accepted time points to 1e-9 (same accept/reject
  sequence);
* evp.power_method on a TFIM chain (tests/golden/evp_power_tfim_d6.npz):
  eigenvalue to 1e-9 relative, eigentensor up to sign;
* TT.norm(p=1) against dense sums.
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


def test_norm_p1(env):
    rng = np.random.default_rng(3)
    x = [np.abs(rng.standard_normal(s)) for s in [(1, 3, 1, 2), (2, 4, 1, 3), (3, 2, 1, 1)]]
    A = [np.abs(rng.standard_normal(s)) for s in [(1, 3, 2, 2), (2, 2, 4, 1)]]
    assert abs(_tt(x).norm(p=1) - tt_ops.matricize(x).sum()) < 1e-12 * tt_ops.matricize(x).sum()
    M = tt_ops.matricize(A)
    assert abs(_tt(A).norm(p=1) - M.sum(axis=0).max()) < 1e-12 * M.sum(axis=0).max()


def test_implicit_euler_vs_reference(env):
    import scikit_tt_amd.solvers.ode as ode
    g = load_golden('ode_signaling_cascade')
    op, iv, ig = _tt(g['tt']('A')), _tt(g['tt']('iv')), _tt(g['tt']('ig'))
    sol = ode.implicit_euler(op, iv, ig, [1.0] * 20, progress=False)
    assert len(sol) == 21
    assert _reldiff(sol[-1].cores[:], g['tt']('ie_last')) < 1e-7
    errs = ode.errors_impl_euler(op, sol, [1.0] * 20)
    np.testing.assert_allclose(errs, g['ie_errors'], rtol=1e-3, atol=1e-12)


def test_trapezoidal_rule_vs_reference(env):
    import scikit_tt_amd.solvers.ode as ode
    g = load_golden('ode_signaling_cascade')
    op, iv, ig = _tt(g['tt']('A')), _tt(g['tt']('iv')), _tt(g['tt']('ig'))
    steps = [0.01] * 5 + [0.5] * 5
    sol = ode.trapezoidal_rule(op, iv, ig, steps, progress=False)
    assert _reldiff(sol[-1].cores[:], g['tt']('tr_last')) < 1e-7
    errs = ode.errors_trapezoidal(op, sol, steps)
    np.testing.assert_allclose(errs, g['tr_errors'], rtol=1e-3, atol=1e-12)


def test_adaptive_step_size_vs_reference(env):
    import scikit_tt_amd.solvers.ode as ode
    g = load_golden('ode_signaling_cascade')
    op, iv, ig = _tt(g['tt']('A')), _tt(g['tt']('iv')), _tt(g['tt']('ig'))
    sol, times = ode.adaptive_step_size(op, iv, ig, 30, step_size_first=1, second_method='two_step_Euler',
                                        progress=False)
    assert len(times) == len(g['ad_times'])
    np.testing.assert_allclose(times, g['ad_times'], rtol=1e-9)
    assert _reldiff(sol[-1].cores[:], g['tt']('ad_last')) < 1e-7


def test_power_method_vs_reference(env):
    import scikit_tt_amd.solvers.evp as evp
    g = load_golden('evp_power_tfim_d6')
    lam, y = evp.power_method(_tt(g['tt']('A')), _tt(g['tt']('x0')), repeats=int(g['repeats']),
                              sigma=float(g['sigma']))
    ref = float(g['eigenvalue'])
    assert abs(lam - ref) <= 1e-9 * abs(ref), (lam, ref)
    yr = g['tt']('x')
    assert min(_reldiff(y.cores[:], yr), _reldiff(tt_ops.scale(y.cores[:], -1.0), yr)) < 1e-6
    # inverse iteration from below the spectrum approaches the ground state
    # from above (a Rayleigh quotient), to the rank-4 limit
    E0 = float(g['exact'])
    assert E0 - 1e-12 <= lam < E0 + 1e-3 * abs(E0)
