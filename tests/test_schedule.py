"""CPU: the launch-time model and the split planner of the mixed sle
schedule (host arithmetic in libttals.so; no device).

* the cost of a system is recomputed here from the shapes: the flops of the
  three-stage local matvec 2(aPa'nb + a'bPnnQ + a'nQbb') summed over the
  positions of one forward and one backward half-sweep, on cases where the
  on-chip/off-chip split is unambiguous (every local problem of at most 96
  unknowns is on-chip, the dense path; one of 32x16x32 never fits 61 KB);
* the planner's modelled launch time is recomputed by an independent list
  schedule of its own split; the split never uses more workgroups than the
  device holds, gives no gang when none shortens the launch, and never
  models a longer launch than one workgroup per system.
"""
import heapq

import numpy as np
import pytest

from scikit_tt_amd import tune, workloads as W

DEFAULT = np.array([1.0e-7, 3.4e-7, 0.1, 1.0, 0.65, 0.40, 0.36, 0.32, 0.32])


def _shapes(d, n, r, R):
    A = [np.zeros((1 if i == 0 else R, n, n, 1 if i == d - 1 else R)) for i in range(d)]
    x = [np.zeros((1 if i == 0 else r, n, 1, 1 if i == d - 1 else r)) for i in range(d)]
    return A, x


def _cost_here(A, x):
    d = len(A)
    tot = 0.0
    for i in range(d):
        a, n, b = x[i].shape[0], x[i].shape[1], x[i].shape[3]
        P, Q = A[i].shape[0], A[i].shape[3]
        mv = 2.0 * (a * P * a * n * b + a * b * P * n * n * Q + a * n * Q * b * b)
        tot += (1 if i == d - 1 else 2) * mv
    return tot


@pytest.mark.parametrize('d,n,r,R,onchip', [(5, 4, 2, 2, True), (8, 2, 4, 3, True), (3, 16, 32, 3, False)])
def test_cost_matches_formula(d, n, r, R, onchip):
    A, x = _shapes(d, n, r, R)
    c = tune.costs_of_shapes([A], [x])[0]
    ref = _cost_here(A, x)
    if onchip:
        assert c[1] == 0.0 and c[0] == pytest.approx(ref, rel=1e-15)
    else:
        # the outer positions (a or b = 1) are small enough for the arena
        assert c[1] > 0.5 * ref and c[0] + c[1] == pytest.approx(ref, rel=1e-15)
    assert c[2] == 2 * d - 1


def _lpt(t1, tg, G, slots):
    heap = []
    end = 0.0
    for s in range(len(G)):
        if G[s] > 0:
            heap += [tg[s]] * G[s]
            end = max(end, tg[s])
    heap += [0.0] * (slots - len(heap))
    heapq.heapify(heap)
    for s in sorted(range(len(t1)), key=lambda k: -t1[k]):
        if G[s] > 0:
            continue
        e = heapq.heappop(heap) + t1[s]
        end = max(end, e)
        heapq.heappush(heap, e)
    return end


def _model(ops, xs, repeats=1, p=DEFAULT):
    """One-workgroup times and gang times as functions of the gang factor."""
    c = tune.costs_of_shapes(ops, xs)
    t1 = tune.predict(c, repeats, p)
    return t1, lambda f: repeats * (p[1] * (c[:, 0] + c[:, 1]) * f + p[2] * c[:, 2])


def test_plan_of_the_uneven_batch():
    inst = [W.shifted_batch_instance(s) for s in range(512)]
    ops, xs = [i[0] for i in inst], [i[1] for i in inst]
    g, ms = tune.plan_of_shapes(ops, xs, repeats=1, slots=256)
    t1, tgf = _model(ops, xs)
    assert g.sum() <= 256 and set(np.unique(g)) <= {0, 2, 4, 8, 16, 32}
    k = np.log2(np.maximum(g, 1)).astype(int)
    tg = tgf(DEFAULT[3 + k])
    assert ms == pytest.approx(_lpt(t1, tg, g, 256), rel=1e-12)
    fused = _lpt(t1, tg, np.zeros_like(g), 256)
    assert ms < 0.5 * fused            # the largest systems set the one-workgroup launch
    assert g[np.argmax(t1)] >= 8       # and the largest gets a large gang
    # gangs go to the slowest systems: every ganged system is slower alone than every single one
    assert t1[g > 0].min() >= t1[g == 0].max()


def test_plan_without_a_dominant_system_has_no_gangs():
    ops, xs = [], []
    for s in range(64):
        A, x = _shapes(6, 4, 3, 2)
        ops.append(A)
        xs.append(x)
    g, ms = tune.plan_of_shapes(ops, xs, repeats=2, slots=256)
    assert (g == 0).all()
    t1, _ = _model(ops, xs, repeats=2)
    assert ms == pytest.approx(t1.max(), rel=1e-12)


def test_plan_respects_the_slots():
    inst = [W.shifted_batch_instance(s) for s in range(64)]
    ops, xs = [i[0] for i in inst], [i[1] for i in inst]
    for slots in (1, 8, 40):
        g, ms = tune.plan_of_shapes(ops, xs, repeats=1, slots=slots)
        assert g.sum() <= slots
        t1, tgf = _model(ops, xs)
        k = np.log2(np.maximum(g, 1)).astype(int)
        assert ms == pytest.approx(_lpt(t1, tgf(DEFAULT[3 + k]), g, slots), rel=1e-12)
        assert ms <= _lpt(t1, t1, np.zeros_like(g), slots) * (1 + 1e-12)
