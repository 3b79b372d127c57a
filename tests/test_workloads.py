"""Host-side workload helpers (no GPU)."""
import numpy as np

from oracle import tt_ops
from scikit_tt_amd import workloads as W


def test_right_orthonormal_same_tensor_unit_norm():
    _, y0, _, _ = W.tfim_chain(seed=3, d=10, rank=6)
    x = W.right_orthonormal(y0)
    a, b = tt_ops.matricize(y0), tt_ops.matricize(x)
    np.testing.assert_allclose(b, a / np.linalg.norm(a), atol=1e-14)
    assert abs(np.linalg.norm(b) - 1.0) < 1e-14
    for i in range(1, len(x)):
        r = x[i].shape[0]
        m = x[i].reshape(r, -1)
        np.testing.assert_allclose(m @ m.T, np.eye(r), atol=1e-13)
    # ranks shrink only where the unfolding is short (n r' < r)
    assert [c.shape[0] for c in x][1:] == [min(y0[i].shape[0], 2 * x[i].shape[3]) for i in range(1, len(x))]


def test_right_orthonormal_without_normalisation_keeps_the_tensor():
    _, y0, _, _ = W.tfim_chain(seed=4, d=6, rank=4)
    x = W.right_orthonormal(y0, normalize=False)
    np.testing.assert_allclose(tt_ops.matricize(x), tt_ops.matricize(y0), rtol=1e-13, atol=1e-13)
