"""Benchmark: TT-ALS solves/s on one MI355X.

Workload (one "step" = one pass of the hot path over one batch): a batch of
B (default 2048) independent transverse-field Ising chains, d = 32 sites, n = 2, MPO rank
3, couplings J_i, g_i ~ U[0.5, 1.5] (seed = instance index), random N(0,1)
initial TT of rank 16 right-orthonormalised and normalised, evp.als lowest
eigenpair ('eig', sigma = -100), 6 sweeps, solved in ONE engine launch.

Printed (one JSON line): batched solves/s (value), ms per step, the plain
drop-in path (one system per call) solves/s, eigenvalue agreement with the
per-system FP64 result and with the free-fermion closed form, per-step
latency p50/p99, the roofline of the engine kernel (algorithmic FP64 flops
counted from shapes and Krylov iteration counts / launch time measured with
events on the launch stream) and a CPU baseline: the numpy oracle timed on
one sweep of one instance on this host.
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


def build_batch(B, d, rank, seed0=0):
    """B TFIM instances; operators and right-orthonormal unit-norm initial
    guesses uploaded with one host-to-device copy.  Of the host cores only the
    first instance's (operator, raw initial guess) is kept, for the CPU
    baseline: holding every instance's arrays would only feed Python's
    cyclic garbage collector (full collections in the timed loop)."""
    from scikit_tt_amd import workloads as W
    from scikit_tt_amd.tensor_train import from_host_batch
    hops, hx0, exact, first = [], [], [], None
    for s in range(seed0, seed0 + B):
        H, y0, J, g = W.tfim_chain(seed=s, d=d, rank=rank)
        hops.append(H)
        hx0.append(W.right_orthonormal(y0))
        exact.append(W.tfim_ground_energy(J, g))
        if first is None:
            first = (H, y0)
    tts = from_host_batch(hops + hx0)
    return tts[:B], tts[B:], np.array(exact), first


def cpu_baseline(raw_instance, sweeps, sample_sweeps=2):
    """numpy oracle: `sample_sweeps` sweeps of one instance on this host, extrapolated to `sweeps`."""
    from oracle import tt_ops, evp_als
    H, y0 = raw_instance
    y0 = tt_ops.ortho_right(y0)
    y0 = tt_ops.scale(y0, 1.0 / tt_ops.norm(y0))
    try:
        from threadpoolctl import threadpool_info
        threads = max([p.get('num_threads', 1) for p in threadpool_info() if p.get('user_api') == 'blas'] or [1])
    except Exception:  # pragma: no cover
        threads = os.cpu_count() or 1
    t = time.perf_counter()
    evp_als.als(H, y0, repeats=sample_sweeps, conv_eps=0, solver='eig', sigma=-100)
    dt = time.perf_counter() - t
    per_solve = dt * sweeps / sample_sweeps
    return {'value': 1.0 / per_solve, 'unit': 'solves/s', 'cores': int(threads), 'kind': 'port',
            'sample': '%d of %d sweeps of one instance (numpy oracle, dense local eig), extrapolated x%.1f; '
                      '%.2f s measured' % (sample_sweeps, sweeps, sweeps / sample_sweeps, dt)}


def measured_traffic(batch, sites, rank, sweeps, adaptive):
    """HBM bytes per k_als launch for this workload from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, calibrated against a
    known-size kernel), or None if no summary covers it."""
    import re
    path = os.path.join(ROOT, 'profiles', 'k_als_tfim_d%d_r%d_b%d_pmc.md' % (sites, rank, batch))
    if not os.path.exists(path):
        return None, None
    text = open(path).read()
    title = text.splitlines()[0]
    if ('%d sweeps' % sweeps) not in title:
        return None, None
    key = 'adaptive' if adaptive else 'fixed'
    m = re.search(r'^traffic_bytes_per_launch_%s: (\d+)' % key, text, re.M)
    return (int(m.group(1)), os.path.relpath(path, ROOT)) if m else (None, None)


def serving_phase(rate, seconds, seed=7, adaptive=True):
    """Poisson arrivals of mixed evp (TFIM, d 8..32, rank 4..16 log-uniform) and
    sle (-Laplacian, d 6..16, n 4/8, rank 4..8) requests into the
    continuous-batching server; returns latency percentiles and throughput."""
    import torch
    from scikit_tt_amd import workloads as W
    from scikit_tt_amd.tensor_train import from_host_batch
    from scikit_tt_amd.serving import Server
    rng = np.random.default_rng(seed)
    nreq = max(1, int(rate * seconds))
    gaps = rng.exponential(1.0 / rate, nreq)
    reqs = []
    for k in range(nreq):
        if rng.random() < 0.5:
            d = int(rng.integers(8, 33))
            r = int(np.exp(rng.uniform(np.log(4), np.log(16))))
            H, y0, J, g = W.tfim_chain(seed=1000 + k, d=d, rank=r)
            reqs.append(('evp',) + tuple(from_host_batch([H, W.right_orthonormal(y0)])) + (None,))
        else:
            d, n = int(rng.integers(6, 17)), int(rng.choice([4, 8]))
            r = int(rng.integers(4, 9))
            A, x0, b = W.laplace_small(seed=2000 + k, d=d, n=n, rank=r)
            reqs.append(('sle',) + tuple(from_host_batch([A, x0, b])))
    torch.cuda.synchronize()
    srv = Server(slots=4, max_batch=256)
    t0 = time.perf_counter()
    next_t = t0
    for (kind, A, x0, b), gap in zip(reqs, gaps):
        next_t += gap
        while time.perf_counter() < next_t:
            srv.poll()
        if kind == 'evp':
            srv.submit_evp(A, x0, repeats=4, conv_eps=1e-10, solver='eig', sigma=-100.0, adaptive_tol=adaptive)
        else:
            srv.submit_sle(A, x0, b, repeats=2)
        srv.poll()
    srv.drain(timeout=600)
    elapsed = time.perf_counter() - t0
    st = srv.latency_stats()
    bad = sum(1 for t in srv.done if t.info.status not in (0,))
    srv.close()
    return {'rate_req_per_s': rate, 'requests': nreq, 'solves_per_s': round(nreq / elapsed, 2),
            'p50_ms': round(st['p50_ms'], 3), 'p99_ms': round(st['p99_ms'], 3), 'launches': st['launches'],
            'nonzero_status': bad}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--batch', type=int, default=2048)
    ap.add_argument('--sites', type=int, default=32)
    ap.add_argument('--rank', type=int, default=16)
    ap.add_argument('--sweeps', type=int, default=6)
    ap.add_argument('--plain', type=int, default=4, help='systems timed one per call on the plain path')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--fixed-tol', action='store_true',
                    help='time the batch with every local eigenproblem solved to 1e-12 (default: adaptive)')
    ap.add_argument('--serve-rates', default='100,1000', help='Poisson arrival rates (req/s) of the serving phase')
    ap.add_argument('--serve-seconds', type=float, default=1.0)
    ap.add_argument('--copy-bytes', type=int, default=0,
                    help='run one plain device copy of this many bytes before the timed steps (a known-size '
                         'reference for hardware byte counters)')
    args = ap.parse_args()
    if args.plain < 1:
        ap.error('--plain must be at least 1 (the plain drop-in rate is part of the result)')
    if args.gpus != 1:
        print('bench.py measures one GPU only: every box here has one MI355X. Across GPUs the '
              'independent systems of a batch would shard one slice per GPU with no communication.',
              file=sys.stderr)
        sys.exit(2)

    import torch
    from scikit_tt_amd import _lib
    import scikit_tt_amd.solvers.evp as evp
    assert torch.cuda.is_available(), 'bench.py needs a HIP device'
    torch.cuda.set_device(0)
    _lib.load()
    print('[bench] building %d instances' % args.batch, file=sys.stderr, flush=True)
    ops, x0s, exact, first = build_batch(args.batch, args.sites, args.rank)
    kw = dict(repeats=args.sweeps, conv_eps=0.0, solver='eig', sigma=-100.0)
    kw_batch = dict(kw, adaptive_tol=not args.fixed_tol)

    stream = torch.cuda.current_stream()
    if args.copy_bytes > 0:
        # an ordinary elementwise kernel (reads and writes copy_bytes each), not a blit copy
        src = torch.ones(args.copy_bytes // 8, dtype=torch.float64, device='cuda')
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    for w in range(args.warmup):
        evp.als_batch(ops, x0s, **kw_batch)
        print('[bench] warmup %d done' % w, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    lat, kernel_ms, call_ms, flops, lams = [], [], [], 0.0, None
    ctx = _lib.context(0)
    ms = ctypes.c_double()
    t0 = time.perf_counter()
    for k in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ts = time.perf_counter()
        e0.record(stream)
        out, infos = evp.als_batch(ops, x0s, return_info=True, **kw_batch)
        e1.record(stream)
        torch.cuda.synchronize()
        lat.append(1e3 * (time.perf_counter() - ts))
        call_ms.append(e0.elapsed_time(e1))
        # the engine kernel alone: events recorded by the library right around its launch
        _lib.check(_lib.load().ttals_last_launch_ms(ctx, ctypes.byref(ms)), ctx, 'last_launch_ms')
        kernel_ms.append(ms.value)
        flops = sum(i.flops for i in infos)
        lams = np.array([o[0] for o in out])
        bad = [i.status for i in infos if i.status not in (0,)]
        print('[bench] step %d: %.1f ms, status %s' % (k, lat[-1], 'ok' if not bad else bad), file=sys.stderr,
              flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = args.batch * args.steps / elapsed

    # FP64 reference for every result: the same batch with every local
    # eigenproblem solved to 1e-12 (what the plain drop-in computes)
    torch.cuda.synchronize()
    tf = time.perf_counter()
    ref_out = evp.als_batch(ops, x0s, **kw)
    torch.cuda.synchronize()
    fixed_rate = args.batch / (time.perf_counter() - tf)
    ref = np.array([o[0] for o in ref_out])
    err_vs_fp64 = np.abs(lams - ref)

    # plain drop-in path: one system per call
    P = min(args.plain, args.batch)
    torch.cuda.synchronize()
    tp = time.perf_counter()
    plain = [evp.als(ops[s], x0s[s], **kw)[0] for s in range(P)]
    torch.cuda.synchronize()
    plain_rate = P / (time.perf_counter() - tp)
    plain_vs_batch_fixed = float(np.max(np.abs(ref[:P] - np.array(plain))))
    err_exact = lams - exact
    print('[bench] plain path: %.2f solves/s' % plain_rate, file=sys.stderr, flush=True)

    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    traffic, traffic_src = measured_traffic(args.batch, args.sites, args.rank, args.sweeps, not args.fixed_tol)
    achieved = flops / avg_kernel_s / 1e12
    line = {
        'metric': 'TT-ALS solves/s at rel.res <=1e-8 (eigenvalue error <=1e-10), batched vs plain; p50/p99 latency',
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
        'config': {'workload': 'evp.als lowest eigenpair of %d TFIM chains (d=%d, n=2, MPO rank 3, TT rank %d, '
                               '%d sweeps) in one launch' % (args.batch, args.sites, args.rank, args.sweeps),
                   'batch': args.batch, 'sites': args.sites, 'rank': args.rank, 'sweeps': args.sweeps,
                   'parallelism': 'one workgroup per system'},
        'local_tol': 'fixed 1e-12' if args.fixed_tol else 'adaptive (inexact ALS: min(1e-9, 0.1*sqrt(dlambda/lambda)))',
        'plain_fp64_solves_per_s': round(plain_rate, 3),
        'batched_fixed_tol_solves_per_s': round(fixed_rate, 3),
        'speedup_batched_vs_plain': round(value / plain_rate, 2),
        'eig_err_vs_fp64_max': float(err_vs_fp64.max()),
        'eig_err_vs_fp64_median': float(np.median(err_vs_fp64)),
        'eig_within_1e-10': int(np.sum(err_vs_fp64 <= 1e-10)),
        'plain_vs_batched_fixed_max': plain_vs_batch_fixed,
        'eig_err_vs_closed_form_max': float(np.max(np.abs(err_exact))),
        'latency_ms_p50': round(float(np.percentile(lat, 50)), 3),
        'latency_ms_p99': round(float(np.percentile(lat, 99)), 3),
        'roofline': {'bound': 'mfma', 'achieved': achieved, 'peak': PEAK_FP64_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': achieved / PEAK_FP64_TFLOPS, 'traffic': traffic, 'traffic_source': traffic_src,
                     'kernel': 'k_als<double>', 'flops_per_launch': flops, 'launch_ms': avg_kernel_s * 1e3,
                     'call_ms': float(np.mean(call_ms))},
    }
    if args.serve_rates:
        line['serving'] = []
        for rate in [float(r) for r in args.serve_rates.split(',') if r]:
            print('[bench] serving at %.0f req/s' % rate, file=sys.stderr, flush=True)
            line['serving'].append(serving_phase(rate, args.serve_seconds, adaptive=not args.fixed_tol))
    if not args.no_cpu:
        print('[bench] cpu baseline', file=sys.stderr, flush=True)
        line['cpu_baseline'] = cpu_baseline(first, args.sweeps)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
