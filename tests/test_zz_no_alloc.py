"""GPU: the allocation-free entry point ttals_sle_als_batch_static.

It must give exactly what ttals_sle_als_batch gives (same kernels, same
inputs: bit-identical cores, ranks and reports), on two calls in a row --
the second with a smaller batch, so the static scratch still holds the first
call's larger contents -- and reject a batch larger than its static arrays.
"""
import ctypes

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
    lib = _lib.load()
    return lib, _lib.context(0), _lib


def _systems(count, seed0):
    out = []
    for s in range(count):
        A, x, b = W.shifted_batch_instance(seed0 + s, d_range=(4, 9), n_choices=(2, 4, 8), r_range=(2, 12))
        out.append((A, x, b))
    return out


def _run(env, systems, entry):
    """Solve through the C ABI with `entry`; returns (reports, [cores per system], [ranks per system])."""
    lib, ctx, _lib = env
    from scikit_tt_amd.tensor_train import from_host_batch
    from scikit_tt_amd.solvers import _engine
    n = len(systems)
    tts = from_host_batch([A for A, _, _ in systems] + [x for _, x, _ in systems] + [b for _, _, b in systems])
    As, Xs, Bs = tts[:n], tts[n:2 * n], tts[2 * n:]
    dA, pA = _engine.descs(As)
    dX, pX = _engine.descs(Xs)
    dB, pB = _engine.descs(Bs)
    (buf, slots), dO, pO = _engine.outputs(Xs)
    buf.fill_(np.nan)  # the engine must write every output element it reports
    res = (_lib.Result * n)()
    rc = getattr(lib, entry)(ctx, None, n, pA, pX, pB, pO, 2, 0, 1e-12, 4000, res)
    torch.cuda.synchronize()
    assert rc == 0, (entry, rc, lib.ttals_last_error(ctx))
    reports = np.frombuffer(res, dtype=_engine.RESULT_DTYPE).copy()
    cores, ranks = [], []
    host = buf.cpu().numpy()
    for s in range(n):
        r = dO.ranks_of(s)
        ranks.append(r)
        nd = [x.shape[1] for x in systems[s][1]]
        cores.append([host[int(slots[s][i]):int(slots[s][i]) + r[i] * nd[i] * r[i + 1]].copy()
                      for i in range(len(nd))])
    return reports, cores, ranks


def _same(a, b):
    ra, ca, ka = a
    rb, cb, kb = b
    assert ka == kb
    for x, y in zip(ca, cb):
        for u, v in zip(x, y):
            np.testing.assert_array_equal(u, v)
    for f in ('status', 'sweeps', 'krylov_iters', 'local_solves', 'worst_residual', 'flops', 'bytes'):
        np.testing.assert_array_equal(ra[f], rb[f])


def test_static_entry_equals_allocating_entry(env):
    big = _systems(24, 100)
    small = _systems(5, 300)
    first = _run(env, big, 'ttals_sle_als_batch_static')
    _same(first, _run(env, big, 'ttals_sle_als_batch'))
    second = _run(env, small, 'ttals_sle_als_batch_static')   # smaller: the scratch holds the first call's data
    _same(second, _run(env, small, 'ttals_sle_als_batch'))
    assert (first[0]['status'] == 0).all() and (second[0]['status'] == 0).all()


def test_static_entry_rejects_oversized_batch(env):
    lib, ctx, _lib = env
    from scikit_tt_amd.tensor_train import from_host_batch
    from scikit_tt_amd.solvers import _engine
    # one system whose Krylov basis alone (64 x N doubles, N = 96*16*96) exceeds the 64 MiB scratch
    A, x, b = W.shifted_batch_instance(0, d_range=(4, 4), n_choices=(16,), r_range=(96, 96))
    As, Xs, Bs = from_host_batch([A, x, b])
    dA, pA = _engine.descs([As])
    dX, pX = _engine.descs([Xs])
    dB, pB = _engine.descs([Bs])
    _, dO, pO = _engine.outputs([Xs])
    res = (_lib.Result * 1)()
    rc = lib.ttals_sle_als_batch_static(ctx, None, 1, pA, pX, pB, pO, 1, 0, 1e-12, 100, res)
    assert rc == _lib.E_ARG
    assert b'device scratch' in lib.ttals_last_error(ctx)
