"""Launch-time model of the mixed sle schedule, fitted by measurement.

A batched launch lasts as long as its slowest system.  With
``sle.als_batch(..., schedule='auto')`` the library gives the systems that
would set that time a gang of workgroups (the multi-workgroup engine of
``sle.als_gang``) and runs the rest one workgroup each, all in one launch.
Which systems, and how large a gang each gets, follows from a model of the
time of one system:

    t1 = repeats * (on * F_on + off * F_off + solve * n_solve)     [ms]

where F_on / F_off are the flops of the local matvec summed over one sweep
at the positions whose local problem fits / does not fit the on-chip arena
(``costs``), and a gang of 2^k workgroups takes

    repeats * (off * (F_on + F_off) * gang[k] + solve * n_solve)

(a gang keeps its operands in HBM/L2 and splits the flops, not the
per-solve latency).

``calibrate`` times single systems and gangs of the batch it is given on the
device, fits the nine parameters and installs them on the device's context;
``get_model`` / ``set_model`` read and write them (e.g. to reuse a fit).
There is no counterpart in the reference, which solves one system at a time.
"""
import ctypes

import numpy as np

from . import _lib

PARAMS = ('on', 'off', 'solve', 'gang1', 'gang2', 'gang4', 'gang8', 'gang16', 'gang32')
GANGS = (1, 2, 4, 8, 16, 32)


def costs(operators, initial_guesses, device=None):
    """(count, 3) array: on-chip matvec flops, off-chip matvec flops and local
    solves of one sweep for each system (host arithmetic on the shapes)."""
    from .solvers import _engine
    count = len(operators)
    if count == 0 or len(initial_guesses) != count:
        raise ValueError('operators and initial guesses must be non-empty lists of equal length')
    _, A = _engine.descs(operators)
    _, X = _engine.descs(initial_guesses)
    out = np.zeros((count, 3))
    dev = operators[0].device.index if device is None else device
    ctx = _lib.context(dev)
    _lib.check(_lib.load().ttals_sle_cost(ctx, count, A, X, out.ctypes.data), ctx, 'sle_cost')
    return out


def _shape_descs(trains):
    """ttals_tt[] holding only the shapes of host trains (lists of 4-d cores);
    the core pointers are placeholders that ttals_sle_cost never reads."""
    orders, rows, cols, ranks = [], [], [], []
    for t in trains:
        orders.append(len(t))
        rows += [c.shape[1] for c in t]
        cols += [c.shape[2] for c in t]
        ranks += [c.shape[0] for c in t] + [t[-1].shape[3]]
    return _lib.DescBatch(orders, rows, cols, ranks, np.ones(sum(orders), dtype=np.uint64))


def costs_of_shapes(operators, initial_guesses):
    """``costs`` for host trains (lists of 4-d numpy cores), on the CPU with
    the default arena: no device or context is needed."""
    if len(operators) == 0 or len(initial_guesses) != len(operators):
        raise ValueError('operators and initial guesses must be non-empty lists of equal length')
    dA, dX = _shape_descs(operators), _shape_descs(initial_guesses)
    out = np.zeros((len(operators), 3))
    rc = _lib.load().ttals_sle_cost(None, len(operators), dA.ptr, dX.ptr, out.ctypes.data)
    if rc != 0:
        raise ValueError('ttals_sle_cost failed with %d (bad shapes)' % rc)
    return out


def plan_of_shapes(operators, initial_guesses, repeats=1, slots=256):
    """The split the mixed schedule makes of host trains under the default
    model: (gang size per system, 0 = one workgroup; modelled launch ms)."""
    dA, dX = _shape_descs(operators), _shape_descs(initial_guesses)
    gangs = np.zeros(len(operators), dtype=np.int32)
    ms = ctypes.c_double()
    rc = _lib.load().ttals_sle_plan(None, len(operators), dA.ptr, dX.ptr, int(repeats), int(slots),
                                    gangs.ctypes.data, ctypes.byref(ms))
    if rc != 0:
        raise ValueError('ttals_sle_plan failed with %d' % rc)
    return gangs, ms.value


def get_model(device=0):
    """The context's nine model parameters (see PARAMS)."""
    ctx = _lib.context(device)
    p = np.zeros(9)
    _lib.check(_lib.load().ttals_get_cost_model(ctx, p.ctypes.data, 9), ctx, 'get_cost_model')
    return p


def set_model(params, device=0):
    p = np.ascontiguousarray(params, dtype=np.float64)
    if p.shape != (9,):
        raise ValueError('the model has 9 parameters %s' % (PARAMS,))
    ctx = _lib.context(device)
    _lib.check(_lib.load().ttals_set_cost_model(ctx, p.ctypes.data, 9), ctx, 'set_cost_model')


def predict(c, repeats, params):
    """Modelled one-workgroup time (ms) of systems with costs ``c``."""
    on, off, solve = params[:3]
    return repeats * (on * c[:, 0] + off * c[:, 1] + solve * c[:, 2])


def _kernel_ms(ctx):
    ms = ctypes.c_double()
    _lib.check(_lib.load().ttals_last_launch_ms(ctx, ctypes.byref(ms)), ctx, 'last_launch_ms')
    return ms.value


def _fit_rates(C, t, repeats, fallback):
    """on, off, solve rates: least squares of repeats * C @ rates = t in
    relative error, rates >= 0 (a rate with no cost behind it keeps its
    fallback value)."""
    M = repeats * C / t[:, None]
    try:
        from scipy.optimize import nnls
        rates, _ = nnls(M, np.ones(len(t)))
    except ImportError:
        rates = np.linalg.lstsq(M, np.ones(len(t)), rcond=None)[0]
    rates = np.maximum(rates, 1e-15)
    for j in range(3):
        if not np.any(C[:, j] > 0):
            rates[j] = fallback[j]
    return rates


def calibrate(operators, initial_guesses, right_hand_sides, repeats=1, sample=10, gang_sample=3, gangs=GANGS[1:],
              precision='fp64', install=True, load_passes=3):
    """Fit the model on the device from timings of systems of this batch.

    ``sample`` systems spread over the batch's range of modelled cost (the
    largest always included) are timed alone on one workgroup; the on, off
    and solve rates are their least-squares fit in relative error.  The
    ``gang_sample`` largest are timed on gangs of each size in ``gangs``
    (at most their widest rank); gang[k] is the median over them of
    (gang time - modelled per-solve time) / off-chip-priced matvec time.

    A system runs slower inside a full launch than alone (the systems share
    L2 and HBM), and not uniformly: large global-memory systems more than
    gangs and small on-chip systems.  ``load_passes`` times the batch itself
    that many times with the current model and, from every system's device
    time inside the launch (``device_ms`` of its report), scales the
    off-chip rate by the median slowdown of the large single-workgroup
    systems and the gang factors by the gangs' slowdown relative to theirs.
    Returns a dict with the fitted parameters, the timings and each load
    pass's planned and measured launch; ``install`` sets them on the context
    (load passes need it).
    """
    import scikit_tt_amd.solvers.sle as sle
    count = len(operators)
    dev = operators[0].device.index
    ctx = _lib.context(dev)
    c = costs(operators, initial_guesses)
    p0 = get_model(dev)
    t_model = predict(c, repeats, p0)
    order = np.argsort(-t_model, kind='stable')
    k = min(sample, count)
    # the largest systems plus a log-spread over the rest
    picks = list(order[:max(1, k // 3)])
    rest = order[len(picks):]
    if len(rest) and k > len(picks):
        idx = np.unique(np.round(np.geomspace(1, len(rest), k - len(picks))).astype(int) - 1)
        picks += [int(rest[i]) for i in idx]
    picks = [int(s) for s in dict.fromkeys(picks)]
    t1 = []
    for s in picks:
        best = np.inf
        for _ in range(2):
            sle.als_batch([operators[s]], [initial_guesses[s]], [right_hand_sides[s]], repeats=repeats)
            best = min(best, _kernel_ms(ctx))
        t1.append(best)
    t1 = np.array(t1)
    C = c[picks]
    rates = _fit_rates(C, t1, repeats, p0)
    params = p0.copy()
    params[:3] = rates
    # gangs on the largest systems
    big = [int(s) for s in order[:min(gang_sample, count)]]
    ratios = {G: [] for G in gangs}
    gang_ms = {}
    for s in big:
        tmv = repeats * rates[1] * (c[s, 0] + c[s, 1])
        tov = repeats * rates[2] * c[s, 2]
        wmax = max(initial_guesses[s].ranks)
        for G in gangs:
            if G > min(32, wmax):
                continue
            sle.als_gang([operators[s]], [initial_guesses[s]], [right_hand_sides[s]], repeats=repeats,
                         precision=precision, gang=G)
            ms = _kernel_ms(ctx)
            gang_ms[(s, G)] = ms
            ratios[G].append(max(ms - tov, 1e-3 * ms) / tmv)
    for j, G in enumerate(GANGS):
        if j == 0:
            params[3] = 1.0
        elif G in ratios and ratios[G]:
            params[3 + j] = float(np.median(ratios[G]))
        else:  # not measured: no gain assumed over the next smaller gang
            params[3 + j] = params[2 + j]
    if install:
        set_model(params, dev)
    fit = predict(C, repeats, params)
    out = {'params_alone': dict(zip(PARAMS, params.tolist())),
           'systems_timed': len(picks),
           'one_workgroup_ms': t1.tolist(),
           'fit_rel_err_max': float(np.max(np.abs(fit - t1) / t1)),
           'gang_ms': {'%d/G%d' % key: v for key, v in gang_ms.items()},
           'load_passes': []}
    if install:
        lib = _lib.load()
        for _ in range(int(load_passes)):
            g = np.zeros(count, dtype=np.int32)
            _, infos = sle.als_batch(operators, initial_guesses, right_hand_sides, repeats=repeats, schedule='auto',
                                     gang_precision=precision, return_info=True)
            launch = _kernel_ms(ctx)
            planned = ctypes.c_double()
            _lib.check(lib.ttals_last_plan_ms(ctx, ctypes.byref(planned)), ctx, 'last_plan_ms')
            _lib.check(lib.ttals_last_plan(ctx, g.ctypes.data, count), ctx, 'last_plan')
            t = infos.array['device_ms'].copy()
            # in-launch over modelled time: the large single-workgroup systems
            # (global-memory bound, the ones a plan leaves ungated) and the gangs
            t_model = predict(c, repeats, params)
            tmv = repeats * params[1] * (c[:, 0] + c[:, 1])
            tov = repeats * params[2] * c[:, 2]
            gang_model = np.zeros(count)
            for j, G in enumerate(GANGS):
                sel = g == G
                if j > 0 and sel.any():
                    gang_model[sel] = tmv[sel] * params[3 + j] + tov[sel]
            big = (g == 0) & (t > 0) & (t_model >= 0.3 * t_model.max())
            ganged = (g > 0) & (t > 0) & (gang_model > 0)
            if big.sum() >= 3:
                f_s = float(np.median(t[big] / t_model[big]))
                f_g = float(np.median(t[ganged] / gang_model[ganged])) if ganged.any() else f_s
                params[1] *= f_s          # off-chip rate of single systems under load
                params[4:] *= f_g / f_s   # gangs keep their own slowdown
            set_model(params, dev)
            out['load_passes'].append({'planned_ms': planned.value, 'launch_ms': launch,
                                       'gang_sizes': {int(k): int(v) for k, v in zip(*np.unique(g, return_counts=True))}})
    out['params'] = dict(zip(PARAMS, params.tolist()))
    return out

