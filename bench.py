"""Benchmark: TT-ALS solves/s on one MI355X at rel.res <= 1e-8.

Headline workload (one "step" = one pass of the hot path over one batch):
the uneven batch of 512 independent sle.als systems -- instance s drawn
with rng(s): d ~ U{8..24}, n in {4, 8, 16}, solution rank r ~ U{4..48};
A = d*I + a rank-3 TT perturbation with N(0, 0.1^2) cores (TT rank 4);
b a rank-4 N(0,1) TT; x0 a rank-r N(0,1) TT -- swept `--sweeps` times in
ONE engine launch.  Inputs are resident in HBM before the timed region.

value = solves/s that MEET the gate: after the timed steps every system's
relative residual ||A x - b|| / ||b|| is computed in FP64 on the device (the
reference test's check, tests/test_sle.py:52-53, by a Householder sweep over
the residual train, never a Gram expansion) and only systems with rel.res
<= 1e-8 are counted.  The plain FP64 drop-in (sle.als, one system per call)
is timed on the same systems and gated the same way.

Also printed: the same batch's rel.res distribution and rank-floor check
(one more sweep changes nothing for systems whose rank cannot hold a 1e-8
solution), p50/p99 step latency, the roofline of the engine kernel
(algorithmic FP64 flops counted by the engine from shapes and the Krylov
work it did, over the kernel time measured with HIP events on the launch
stream; beside it the algorithmic bytes of the local matvecs, the stacks,
operator core and iterate read once per matvec plus the input and output
vectors where they live in global memory, against the measured traffic), a
secondary evp workload (2048 TFIM chains, d=32, rank 16, 6
sweeps) with its eigenvalue error against the reference-written value of
instance 0 and against the FP64 fixed-tolerance run, a serving phase, and
the CPU baseline (the numpy oracle, whole solves, on the subset of the
batch whose dense local problems the CPU can factor).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 matrix peak
PEAK_HBM_GBS = 8000.0
GATE = 1e-8


def log(msg):
    print('[bench] ' + msg, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# workloads

def uneven_instances(B):
    from scikit_tt_amd import workloads as W
    return [W.shifted_batch_instance(s) for s in range(B)]


def upload(instances):
    """Operators, initial guesses and right-hand sides with ONE host-to-device copy."""
    from scikit_tt_amd.tensor_train import from_host_batch
    B = len(instances)
    tts = from_host_batch([A for A, _, _ in instances] + [x for _, x, _ in instances] +
                          [b for _, _, b in instances])
    return tts[:B], tts[B:2 * B], tts[2 * B:]


def local_size_max(inst):
    A, x, b = inst
    r = [c.shape[0] for c in x] + [1]
    n = [c.shape[1] for c in x]
    # ALS keeps min(r n, r') / min(r, n r'): the effective ranks
    d = len(n)
    eff = list(r)
    for i in range(d):
        eff[i + 1] = min(eff[i + 1], eff[i] * n[i])
    for i in range(d - 1, -1, -1):
        eff[i] = min(eff[i], n[i] * eff[i + 1])
    return max(eff[i] * n[i] * eff[i + 1] for i in range(d))


# ---------------------------------------------------------------------------
# CPU baseline: the numpy oracle, whole solves

def cpu_baseline(instances, sweeps, budget_s=25.0, nmax=1500):
    """The oracle (dense local LU, numpy/scipy) on the systems of the batch
    whose largest local problem has N <= nmax: 1 warm-up, then >= 5 timed
    whole solves at all host threads and at 1 thread."""
    import scipy
    from threadpoolctl import threadpool_limits, threadpool_info
    from oracle import sle_als
    from scikit_tt_amd import workloads as W
    subset = [k for k, inst in enumerate(instances) if local_size_max(inst) <= nmax]
    # the start right-orthonormalised as the reference's tests do (tests/test_sle.py:35); the device
    # engine does the same internally (ALS is gauge invariant).  Without it the dense local matrices of
    # systems whose initial rank exceeds what a boundary core can hold are singular.
    instances = [(A, W.right_orthonormal(x, normalize=False), b) if k in subset else None
                 for k, (A, x, b) in enumerate(instances)]
    cores = os.cpu_count() or 1
    blas = [p for p in threadpool_info() if p.get('user_api') == 'blas']
    out = {'unit': 'solves/s', 'kind': 'port', 'cores': cores,
           'versions': {'numpy': np.__version__, 'scipy': scipy.__version__,
                        'blas': ['%s %s' % (p.get('internal_api'), p.get('version')) for p in blas]}}
    runs = {}
    for threads in (cores, 1):
        with threadpool_limits(limits=threads, user_api='blas'):
            A, x, b = instances[subset[0]]
            sle_als.als(A, x, b, repeats=sweeps)  # warm-up
            times, used = [], []
            t_all = time.perf_counter()
            for k in subset:
                A, x, b = instances[k]
                t = time.perf_counter()
                sle_als.als(A, x, b, repeats=sweeps)
                times.append(time.perf_counter() - t)
                used.append(k)
                if len(times) >= 5 and time.perf_counter() - t_all > budget_s / 2:
                    break
        times = np.array(times)
        runs[threads] = {'solves': len(times), 'median_s': float(np.median(times)),
                         'p99_s': float(np.percentile(times, 99)), 'solves_per_s': float(len(times) / times.sum()),
                         'systems': used}
    out['value'] = runs[cores]['solves_per_s']
    out['single_thread'] = runs[1]
    out['all_threads'] = runs[cores]
    out['sample'] = ('whole %d-sweep sle.als solves (numpy oracle, dense local LU) of the first systems of the batch '
                     'whose largest local problem has N <= %d (%d of %d systems qualify; larger ones are beyond '
                     'dense CPU factorisation): 1 warm-up + %d timed at %d BLAS threads, %d at 1 thread'
                     % (sweeps, nmax, len(subset), len(instances), runs[cores]['solves'], cores,
                        runs[1]['solves']))
    return out


# ---------------------------------------------------------------------------
# counters committed under profiles/

def measured_traffic(kernel, tag):
    """HBM bytes per launch of `kernel` for workload `tag` from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, calibrated against a
    known-size kernel), or None if no summary covers it."""
    import re
    path = os.path.join(ROOT, 'profiles', '%s_%s_pmc.md' % (kernel, tag))
    if not os.path.exists(path):
        return None, None
    m = re.search(r'^traffic_bytes_per_launch: (\d+)', open(path).read(), re.M)
    return (int(m.group(1)), os.path.relpath(path, ROOT)) if m else (None, None)


def measured_fp64_counters(kernel, tag):
    """MFMA utilisation and executed FP64 flops per launch from the committed
    rocprofv3 FP64 counter summary of `kernel` on workload `tag` ({} if none)."""
    import re
    path = os.path.join(ROOT, 'profiles', '%s_%s_fp64_counters.md' % (kernel, tag))
    if not os.path.exists(path):
        return {}
    text = open(path).read()
    out = {'counters_source': os.path.relpath(path, ROOT)}
    m = re.search(r'^mfma_util_percent: ([0-9.eE+-]+)', text, re.M)
    if m:
        out['mfma_util_percent'] = float(m.group(1))
    m = re.search(r'^fp64_executed_flops_per_launch: (\d+)', text, re.M)
    if m:
        out['fp64_executed_flops_per_launch'] = int(m.group(1))
    return out


# ---------------------------------------------------------------------------
# secondary workload: TFIM chains, evp

def tfim_phase(batch, steps):
    import torch
    from scikit_tt_amd import workloads as W
    from scikit_tt_amd.tensor_train import from_host_batch
    import scikit_tt_amd.solvers.evp as evp
    from tests.conftest import load_golden
    hops, hx0, exact = [], [], []
    for s in range(batch):
        H, y0, J, g = W.tfim_chain(seed=s, d=32, rank=16)
        hops.append(H)
        hx0.append(W.right_orthonormal(y0))
        exact.append(W.tfim_ground_energy(J, g))
    tts = from_host_batch(hops + hx0)
    ops, x0s = tts[:batch], tts[batch:]
    kw = dict(repeats=6, conv_eps=0.0, solver='eig', sigma=-100.0)
    evp.als_batch(ops, x0s, adaptive_tol=True, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = evp.als_batch(ops, x0s, adaptive_tol=True, **kw)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    lams = out.eigenvalues
    ref = evp.als_batch(ops, x0s, **kw).eigenvalues  # every local eigenproblem to 1e-12
    gold = float(load_golden('evp_tfim_d32_c2')['eigenvalue'])  # written by the reference (instance 0)
    return {'workload': 'evp.als lowest eigenpair of %d TFIM chains (d=32, n=2, MPO rank 3, TT rank 16, 6 sweeps, '
                        "'eig', sigma=-100), adaptive local tolerance, one launch" % batch,
            'solves_per_s': round(batch / dt, 2), 'ms_per_step': round(1e3 * dt, 3),
            'eig_err_vs_reference_instance0': abs(float(lams[0]) - gold),
            'eig_err_vs_fp64_fixed_tol_max': float(np.max(np.abs(lams - ref))),
            'eig_within_1e-10_of_fp64': int(np.sum(np.abs(lams - ref) <= 1e-10)),
            'eig_err_vs_closed_form_max': float(np.max(np.abs(lams - np.array(exact))))}


# ---------------------------------------------------------------------------
# serving

def serving_pool(size, seed=700000):
    """Requests of the serving workload (scikit_tt_amd.workloads.serving_request:
    sle or evp, d ~ U{8..64}, n in {2,4,8,16}, ranks log-uniform in [8,128]),
    uploaded once; arrivals draw from this pool."""
    from scikit_tt_amd import workloads as W
    from scikit_tt_amd.tensor_train import from_host_batch
    pool = []
    for k in range(size):
        q = W.serving_request(seed + k)
        q['dev'] = from_host_batch([q['A'], q['x0'], q['b']] if q['kind'] == 'sle' else [q['A'], q['x0']])
        q['N'] = max(c.shape[0] * c.shape[1] * c.shape[3] for c in q['x0'])
        for key in ('A', 'x0', 'b'):
            q.pop(key, None)
        pool.append(q)
    return pool


LARGE_FROM = 16384  # local size from which a serving request is in the large class (its own slot)


def serving_phase(pool, rate, seconds, seed=7, budget_ms=2000.0, model=None):
    """Poisson arrivals at `rate` req/s for `seconds`, each a request drawn
    from the pool, into the continuous-batching server (4 slots, gangs for
    the large sle systems in FP32 + FP64 refinement, `budget_ms` of device
    time per request; requests with a local size >= LARGE_FROM queue for a
    slot of their own).  Served = solved within the budget AND at the gate:
    sle FP64 rel.res <= 1e-8 (sle.als, 2 sweeps), evp eigenvalue within 1e-10
    of the closed form (evp.als 'eig', sigma = -100, up to 6 sweeps to
    conv_eps 1e-10)."""
    import torch
    from scikit_tt_amd.serving import Server
    from scikit_tt_amd.tensor_train import residual_norms
    rng = np.random.default_rng(seed)
    nreq = max(1, int(round(rate * seconds)))
    gaps = rng.exponential(1.0 / rate, nreq)
    picks = rng.integers(0, len(pool), nreq)
    torch.cuda.synchronize()
    srv = Server(slots=4, max_batch=256, maxit=1000, schedule='auto', gang_precision='mixed', budget_ms=budget_ms,
                 cost_model=model, large_from=LARGE_FROM, large_slots=1)
    tickets = []
    t0 = time.perf_counter()
    next_t = t0
    for gap, k in zip(gaps, picks):
        next_t += gap
        while time.perf_counter() < next_t:
            srv.poll()
        q = pool[k]
        if q['kind'] == 'sle':
            A, x0, b = q['dev']
            tickets.append(srv.submit_sle(A, x0, b, repeats=2))
        else:
            A, x0 = q['dev']
            tickets.append(srv.submit_evp(A, x0, repeats=6, conv_eps=1e-10, solver='eig', sigma=-100.0,
                                          adaptive_tol=True))
        srv.poll()
    srv.drain(timeout=900)
    elapsed = time.perf_counter() - t0
    launches = srv.launches
    # the gate (untimed)
    ok = np.zeros(nreq, dtype=bool)
    sle_k = [i for i, t in enumerate(tickets) if t.served and pool[picks[i]]['kind'] == 'sle']
    if sle_k:
        res, bn = residual_norms([pool[picks[i]]['dev'][0] for i in sle_k], [tickets[i].result for i in sle_k],
                                 [pool[picks[i]]['dev'][2] for i in sle_k])
        ok[sle_k] = res / bn <= GATE
    for i, t in enumerate(tickets):
        q = pool[picks[i]]
        if t.served and q['kind'] == 'evp':
            ok[i] = abs(t.result[0] - q['exact']) <= 1e-10 * abs(q['exact'])
    lat = np.array([t.latency for t in tickets])
    deadline = sum(1 for t in tickets if t.info is not None and t.info.status == 7)
    srv.close()
    out = {'rate_req_per_s': rate, 'requests': nreq, 'served_at_gate': int(ok.sum()),
           'served_at_gate_per_s': round(float(ok.sum()) / elapsed, 2), 'over_budget': deadline,
           'p50_ms': round(float(1e3 * np.percentile(lat, 50)), 1),
           'p99_ms': round(float(1e3 * np.percentile(lat, 99)), 1), 'launches': launches,
           'wall_s': round(elapsed, 2)}
    if ok.any():
        out['served_p50_ms'] = round(float(1e3 * np.percentile(lat[ok], 50)), 1)
        out['served_p99_ms'] = round(float(1e3 * np.percentile(lat[ok], 99)), 1)
    return out


# ---------------------------------------------------------------------------

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--batch', type=int, default=512)
    ap.add_argument('--sweeps', type=int, default=1, help='ALS sweeps (repeats; the reference default is 1)')
    ap.add_argument('--plain', type=int, default=128,
                    help='systems timed one per call on the plain path, spread evenly over the batch '
                         '(0 = every system)')
    ap.add_argument('--tfim-batch', type=int, default=2048, help='secondary evp workload (0 = skip)')
    ap.add_argument('--schedule', default='auto', choices=('auto', 'one'),
                    help="sle batch schedule: 'auto' gangs for the systems that set the launch time (model fitted "
                         "on this batch first), 'one' one workgroup per system")
    ap.add_argument('--model', default='fit',
                    help="launch-time model of the mixed schedule: 'fit' (measured on this batch first) or nine "
                         "comma-separated parameters (scikit_tt_amd.tune.PARAMS), e.g. from an earlier fit")
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--serve-rates', default='10,30,100,300,1000,2000',
                    help='Poisson arrival rates (req/s) of the serving phase (empty: skip)')
    ap.add_argument('--serve-seconds', type=float, default=1.0, help='arrival window per rate')
    ap.add_argument('--serve-pool', type=int, default=64, help='distinct serving requests arrivals draw from')
    ap.add_argument('--serve-budget-ms', type=float, default=2000.0, help='device time budget per request')
    ap.add_argument('--copy-bytes', type=int, default=0,
                    help='run one plain device copy of this many bytes before the timed steps (a known-size '
                         'reference for hardware byte counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: every box here has one MI355X. Across GPUs the '
              'independent systems of a batch would shard one slice per GPU with no communication.',
              file=sys.stderr)
        sys.exit(2)

    import torch
    from scikit_tt_amd import _lib
    import scikit_tt_amd.solvers.sle as sle
    from scikit_tt_amd.tensor_train import residual_norms
    assert torch.cuda.is_available(), 'bench.py needs a HIP device'
    torch.cuda.set_device(0)
    _lib.load()
    log('building %d uneven sle instances' % args.batch)
    instances = uneven_instances(args.batch)
    As, Xs, Bs = upload(instances)
    B = args.batch
    stream = torch.cuda.current_stream()
    if args.copy_bytes > 0:
        # an ordinary elementwise kernel (reads and writes copy_bytes each), not a blit copy
        src = torch.ones(args.copy_bytes // 8, dtype=torch.float64, device='cuda')
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    tuning = None
    if args.schedule == 'auto':
        from scikit_tt_amd import tune
        if args.model == 'fit':
            log('fitting the launch-time model on this batch')
            tt0 = time.perf_counter()
            tuning = tune.calibrate(As, Xs, Bs, repeats=args.sweeps)
            tuning['seconds'] = round(time.perf_counter() - tt0, 2)
        else:
            params = [float(v) for v in args.model.split(',')]
            tune.set_model(params, 0)
            tuning = {'params': dict(zip(tune.PARAMS, params)), 'given': True}
        log('model: %s' % tuning['params'])
    for w in range(args.warmup):
        sle.als_batch(As, Xs, Bs, repeats=args.sweeps, schedule=args.schedule)
        log('warmup %d done' % w)
    torch.cuda.synchronize()
    lat, kernel_ms, flops, sols, infos = [], [], 0.0, None, None
    ctx = _lib.context(0)
    ms = ctypes.c_double()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        sols, infos = sle.als_batch(As, Xs, Bs, repeats=args.sweeps, return_info=True, schedule=args.schedule)
        torch.cuda.synchronize()
        lat.append(1e3 * (time.perf_counter() - ts))
        # the engine kernel alone: events recorded by the library right around its launch
        _lib.check(_lib.load().ttals_last_launch_ms(ctx, ctypes.byref(ms)), ctx, 'last_launch_ms')
        kernel_ms.append(ms.value)
        log('step %d: %.1f ms (kernel %.1f ms)' % (k, lat[-1], kernel_ms[-1]))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    arr = infos.array
    flops = float(arr['flops'].sum())
    mv_bytes = float(arr['bytes'].sum())  # algorithmic bytes of the launch's local matvecs
    status_ok = arr['status'] == 0
    gangs = np.asarray(infos.gangs)
    # the same batch on one workgroup per system, once (reported beside the schedule used)
    one_ms = None
    if args.schedule != 'one':
        sle.als_batch(As, Xs, Bs, repeats=args.sweeps)
        torch.cuda.synchronize()
        _lib.check(_lib.load().ttals_last_launch_ms(ctx, ctypes.byref(ms)), ctx, 'last_launch_ms')
        one_ms = ms.value
        log('one workgroup per system: kernel %.1f ms' % one_ms)

    # the gate: FP64 rel.res of every system (untimed)
    res, bn = residual_norms(As, sols, Bs)
    relres = res / bn
    ok = (relres <= GATE) & status_ok
    n_ok = int(ok.sum())
    value = n_ok * args.steps / elapsed
    # rank floor: one more sweep from the result
    more = sle.als_batch(As, sols, Bs, repeats=1)
    res2, _ = residual_norms(As, more, Bs)
    relres2 = res2 / bn
    fail = ~ok
    nvals = np.array([inst[0][0].shape[1] for inst in instances])
    by_n = {int(n): {'systems': int((nvals == n).sum()), 'meet_gate': int((ok & (nvals == n)).sum())}
            for n in sorted(set(nvals.tolist()))}
    log('gate: %d of %d systems at rel.res <= %.0e' % (n_ok, B, GATE))

    # plain FP64 drop-in: one system per call, same systems, same gate
    P = B if args.plain <= 0 else min(args.plain, B)
    idx = np.linspace(0, B - 1, P).round().astype(int) if P < B else np.arange(B)
    torch.cuda.synchronize()
    tp = time.perf_counter()
    plain_sols = [sle.als(As[s], Xs[s], Bs[s], repeats=args.sweeps) for s in idx]
    torch.cuda.synchronize()
    plain_time = time.perf_counter() - tp
    pres, _ = residual_norms([As[s] for s in idx], plain_sols, [Bs[s] for s in idx])
    prel = pres / bn[idx]
    # the plain path runs k_als; the batch's one-workgroup systems run inside
    # k_hybrid (larger on-chip arena, other QR row grouping) and its ganged
    # systems on gangs: the same iterates to rounding, not bit for bit
    one = gangs[idx] == 0

    def relres_ratio(sel):
        if not sel.any():
            return None
        return float(np.max(np.maximum(prel[sel], 1e-300) / np.maximum(relres[idx][sel], 1e-300)))
    plain_ok = int((prel <= GATE).sum())
    plain_rate = plain_ok / plain_time  # gated solves per second of the sampled systems
    log('plain path: %d systems in %.1f s, %d meet the gate' % (P, plain_time, plain_ok))

    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    kernel = 'k_hybrid' if gangs.sum() > 0 else 'k_als'
    traffic, traffic_src = measured_traffic(kernel, 'uneven_sle_b%d_s%d' % (B, args.sweeps))
    counters = measured_fp64_counters(kernel, 'uneven_sle_b%d_s%d' % (B, args.sweeps))
    achieved = flops / avg_kernel_s / 1e12
    line = {
        'metric': 'TT-ALS solves/s at rel.res <=1e-8, batched vs plain; p50/p99 latency',
        'value': round(value, 3),
        'unit': 'solves/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(1e3 * elapsed / args.steps, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic',
        'config': {'workload': 'sle.als on the uneven batch of %d independent systems (rng(s): d~U{8..24}, '
                               'n in {4,8,16}, rank~U{4..48}; A = d*I + rank-3 N(0,0.1^2) TT perturbation; '
                               'rank-4 N(0,1) rhs; rank-r N(0,1) start), %d sweeps, one launch; only systems at '
                               'FP64 rel.res <= 1e-8 counted' % (B, args.sweeps),
                   'batch': B, 'sweeps': args.sweeps, 'gate': GATE,
                   'parallelism': ('one launch: gangs of workgroups for the systems that set the launch time, one '
                                   'workgroup for the rest' if args.schedule == 'auto' else 'one workgroup per system')},
        'schedule': {'mode': args.schedule, 'gang_sizes': {int(k): int(v) for k, v in enumerate(np.bincount(gangs)) if v},
                     'gang_workgroups': int(gangs.sum()), 'launch_ms': round(avg_kernel_s * 1e3, 3),
                     'one_workgroup_per_system_launch_ms': None if one_ms is None else round(one_ms, 3),
                     'model_fit': tuning},
        'systems_meeting_gate': n_ok,
        'systems_total': B,
        'meet_gate_by_mode_size': by_n,
        'batch_solves_per_s_ungated': round(B * args.steps / elapsed, 3),
        'relres_quantiles': {q: float(np.quantile(relres, float(q))) for q in ('0.1', '0.5', '0.9')},
        'rank_floor_check': {'failing': int(fail.sum()),
                             'failing_improved_2x_by_one_more_sweep': int((relres2[fail] < 0.5 * relres[fail]).sum()),
                             'passing_after_one_more_sweep': int(((relres2 <= GATE) & fail).sum())},
        'nonzero_status': int((~status_ok).sum()),
        'plain_fp64_systems_timed': P,
        'plain_fp64_seconds': round(plain_time, 3),
        'plain_fp64_solves_per_s': round(plain_rate, 3),
        'plain_gate_agrees_with_batched': bool(np.array_equal(prel <= GATE, relres[idx] <= GATE)),
        'plain_vs_batched_one_workgroup_systems_max_relres_ratio': relres_ratio(one),
        'plain_vs_batched_ganged_systems_max_relres_ratio': relres_ratio(~one),
        'speedup_batched_vs_plain': round(value / plain_rate, 2) if plain_rate > 0 else None,
        'latency_ms_p50': round(float(np.percentile(lat, 50)), 3),
        'latency_ms_p99': round(float(np.percentile(lat, 99)), 3),
        'roofline': {'bound': 'mfma', 'achieved': achieved, 'peak': PEAK_FP64_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': achieved / PEAK_FP64_TFLOPS, 'traffic': traffic, 'traffic_source': traffic_src,
                     'kernel': kernel, 'flops_per_launch': flops, 'launch_ms': avg_kernel_s * 1e3,
                     'algorithmic_bytes_per_launch': mv_bytes,
                     'algorithmic_hbm_frac': mv_bytes / avg_kernel_s / (PEAK_HBM_GBS * 1e9),
                     'traffic_over_algorithmic_bytes': (traffic / mv_bytes) if (traffic and mv_bytes > 0) else None,
                     **counters},
    }
    if args.tfim_batch > 0:
        log('tfim evp batch %d' % args.tfim_batch)
        line['tfim_evp'] = tfim_phase(args.tfim_batch, max(1, args.steps))
    if args.serve_rates:
        log('serving: building %d requests' % args.serve_pool)
        pool = serving_pool(args.serve_pool)
        sizes = np.array([q['N'] for q in pool])
        line['serving'] = {'workload': 'Poisson arrivals of mixed sle/evp requests (d ~ U{8..64}, n in {2,4,8,16}, '
                                       'ranks log-uniform [8,128]) drawn from %d distinct requests; 4 slots (one for requests of '
                                       'local size >= 16384), gangs '
                                       'in FP32 + FP64 refinement for large sle systems, %.0f ms device budget per '
                                       'request; served = at rel.res <= 1e-8 (sle) or eigenvalue error <= 1e-10 '
                                       '(evp)' % (args.serve_pool, args.serve_budget_ms),
                           'pool_local_size_quantiles': {q: int(np.quantile(sizes, float(q)))
                                                         for q in ('0.5', '0.9', '0.99')},
                           'rates': []}
        model = None
        if tuning is not None:
            from scikit_tt_amd import tune
            model = [tuning['params'][k] for k in tune.PARAMS]
        for rate in [float(r) for r in args.serve_rates.split(',') if r]:
            log('serving at %.0f req/s' % rate)
            r = serving_phase(pool, rate, args.serve_seconds, budget_ms=args.serve_budget_ms, model=model)
            log('  %s' % r)
            line['serving']['rates'].append(r)
    if not args.no_cpu:
        log('cpu baseline')
        line['cpu_baseline'] = cpu_baseline(instances, args.sweeps)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
