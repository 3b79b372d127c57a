"""Serving loop: every result is the same alone and in the mix (bit-exact), and
requests that arrive while others run are admitted into free slots."""
import numpy as np
import pytest

from oracle import tt_ops
from scikit_tt_amd import workloads as W

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def _mk_requests():
    from scikit_tt_amd.tensor_train import TT
    reqs = []
    for s in range(8):
        if s % 2 == 0:
            H, y0, J, g = W.tfim_chain(seed=100 + s, d=8 + 2 * s, rank=4 + s)
            y0 = tt_ops.ortho_right(y0)
            y0 = tt_ops.scale(y0, 1.0 / tt_ops.norm(y0))
            reqs.append(('evp', TT(H), TT(y0), None))
        else:
            A, x0, b = W.laplace_small(seed=s, d=5 + s, n=4 + (s % 3), rank=3 + (s % 4))
            reqs.append(('sle', TT(A), TT(x0), TT(b)))
    return reqs


def test_alone_equals_mix():
    import scikit_tt_amd.solvers.sle as sle
    import scikit_tt_amd.solvers.evp as evp
    from scikit_tt_amd.serving import Server
    assert torch.cuda.is_available()
    reqs = _mk_requests()
    alone = []
    for kind, A, x0, b in reqs:
        if kind == 'sle':
            alone.append(sle.als(A, x0, b, repeats=2))
        else:
            alone.append(evp.als(A, x0, repeats=3, conv_eps=0.0, solver='eig', sigma=-100.0))
    srv = Server(slots=3, max_batch=3)
    tickets = []
    for k, (kind, A, x0, b) in enumerate(reqs):
        if kind == 'sle':
            tickets.append(srv.submit_sle(A, x0, b, repeats=2))
        else:
            tickets.append(srv.submit_evp(A, x0, repeats=3, conv_eps=0.0, solver='eig', sigma=-100.0))
        srv.poll()  # requests keep arriving while earlier launches run
    srv.drain(timeout=60)
    assert srv.launches >= 3
    for (kind, _, _, _), t, ref in zip(reqs, tickets, alone):
        assert t.done and t.info.status == 0
        if kind == 'sle':
            x, xr = t.result, ref
        else:
            assert t.result[0] == ref[0] and t.result[2] == ref[2]
            x, xr = t.result[1], ref[1]
        assert x.ranks == xr.ranks
        for c1, c2 in zip(x.cores, xr.cores):
            np.testing.assert_array_equal(c1, c2)
    st = srv.latency_stats()
    assert st['completed'] == len(reqs) and st['p99_ms'] >= st['p50_ms'] > 0
    srv.close()


def test_admission_respects_budget():
    from scikit_tt_amd.serving import Server, footprint_bytes
    reqs = _mk_requests()
    kind, A, x0, b = reqs[1]
    need = footprint_bytes(kind, A, x0, b)
    srv = Server(slots=2, max_batch=8, hbm_budget=need + 16)
    t1 = srv.submit_sle(A, x0, b)
    t2 = srv.submit_sle(A, x0, b)
    srv.poll()
    # only one request fits the budget at a time
    assert sum(s.busy for s in srv.slots) == 1 and len(srv.queue) == 1
    srv.drain(timeout=60)
    assert t1.done and t2.done
    with pytest.raises(MemoryError):
        Server(slots=1, hbm_budget=16).submit_sle(A, x0, b)
    srv.close()
