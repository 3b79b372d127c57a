"""Continuous-batching request loop for TT-ALS solves on one GPU.

Requests (``sle.als`` or ``evp.als`` problems whose TTs already live on the
device) are queued on arrival.  The loop keeps ``slots`` engine contexts,
each with its own HIP stream and workspace.  Whenever a slot is free it
admits the waiting requests of one kind/parameter set (FIFO, up to
``max_batch`` and within the HBM budget) as ONE launch.  Launches on
different slots run concurrently on the device, so requests that arrive
while others are on the GPU start as soon as a slot and memory free up
instead of waiting for the running batch.  Cores and stacks stay resident on
the device; only ranks and a few scalars return to the host.

``schedule='one'`` solves every system on one workgroup with a fixed
reduction order, so a result does not depend on which batch or slot it ran
in.  ``schedule='auto'`` gives the sle systems that would set a launch's
time a gang of workgroups (``sle.als_batch(..., schedule='auto')``, in
``gang_precision``); a ganged result depends on its gang size, which the
launch's mix decides, only at the engines' agreement (~1e-9).  A request
that outlives ``budget_ms`` of device time stops before its next local step
with status ``TTALS_S_DEADLINE`` (``Ticket.served`` False).
"""
import ctypes
import collections
import time

import numpy as np
import torch

from . import _lib
from .solvers import _engine
from .solvers import evp as evp_mod
from .tensor_train import TT

__all__ = ['Server', 'footprint_bytes']

_SLE_SOLVERS = {'solve': 0, 'lu': 1}
_EVP_SOLVERS = {'eig': 0, 'eigh': 1}
_KBASIS = 64  # Krylov basis vectors the engine plans per system (Lanczos restart 63, + 1)
_NDENSE = 96


def footprint_bytes(kind, A, x0, b=None, gang_max=0):
    """Device bytes one system needs in a launch: stacks, scratch, Krylov basis, output cores.

    Mirrors the engine's planner (interface stacks on every bond, two
    contraction intermediates of N*max(R, p), the 64 x N Krylov basis, QR and
    LU scratch) plus the output cores; with ``gang_max`` > 0 (an sle request
    that may get a gang of up to that many workgroups) also the gang's
    per-member intermediates, its vectors and FP32 copies.
    """
    d = x0.order
    r, R = x0.ranks, A.ranks
    p = b.ranks if b is not None else [1] * (d + 1)
    n = x0.row_dims
    stacks = sum(r[i] * R[i] * r[i] + p[i] * r[i] for i in range(d + 1))
    N = [r[i] * n[i] * r[i + 1] for i in range(d)]
    nmax = max(N)
    tmax = max(N[i] * max(R[i], R[i + 1], p[i], p[i + 1]) for i in range(d))
    rmax = max(r)
    scratch = 4 * nmax + 2 * tmax + _KBASIS * nmax + _NDENSE * (_NDENSE + 1) + rmax * (rmax + 1)
    cores = sum(N)
    best = cores if kind == 'evp' else 0
    gang = 0
    if gang_max > 0 and kind == 'sle':
        # t1/t2 per member, r/p/q/xt and the (41 + 1) x N GMRES basis in FP64 and FP32
        gang = 2 * gang_max * tmax + 6 * nmax + (3 * 42 * nmax) // 2 + gang_max * (132 + 4 * rmax * rmax)
    return 8 * (2 * stacks + scratch + 2 * cores + best + gang) + 4096


class Ticket:
    """Handle of one request; ``result`` is set when it has been solved."""

    __slots__ = ('rid', 'kind', 'key', 'A', 'x0', 'b', 'bytes', 'cls', 't_arrival', 't_start', 't_done', 'result',
                 'info')

    def __init__(self, rid, kind, key, A, x0, b, nbytes, cls=0):
        self.rid, self.kind, self.key, self.cls = rid, kind, key, cls
        self.A, self.x0, self.b = A, x0, b
        self.bytes = nbytes
        self.t_arrival = time.perf_counter()
        self.t_start = self.t_done = None
        self.result = self.info = None

    @property
    def done(self):
        return self.t_done is not None

    @property
    def served(self):
        """Solved within the budget with no failure status (NOCONV counts as served)."""
        return self.info is not None and self.info.status in (0, _lib.S_NOCONV)

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


class _Slot:
    def __init__(self, device, cls=0):
        self.cls = cls  # size class of the requests it admits
        self.ctx = _lib.new_context(device)
        self.stream = torch.cuda.Stream(device=device)  # default priority
        self.batch = None  # (tickets, out cores, out packs, descriptor arrays)

    @property
    def busy(self):
        return self.batch is not None


class Server:
    """Continuous-batching TT-ALS server on one device.

    slots: concurrent launches in flight; max_batch: systems per launch;
    hbm_budget: bytes the in-flight launches may occupy (default: 80% of the
    free device memory at start-up).  ``large_from`` (local size, i.e. the
    largest r n r' of a request's start) splits requests into two size
    classes with their own queues order and ``large_slots`` of the slots
    reserved for the large class, so that small requests are not held behind
    a launch that contains a large one (None: one class).
    """

    def __init__(self, device=0, slots=4, max_batch=128, hbm_budget=None, tol=1e-12, maxit=4000, schedule='one',
                 gang_precision='fp64', budget_ms=None, cost_model=None, large_from=None, large_slots=1):
        if not torch.cuda.is_available():
            raise RuntimeError('Server needs a HIP device')
        if schedule not in ('one', 'auto'):
            raise ValueError("schedule must be 'one' or 'auto'")
        if gang_precision not in ('fp64', 'mixed'):
            raise ValueError("gang_precision must be 'fp64' or 'mixed'")
        self.device = device
        self.lib = _lib.load()
        self.gang_max = 32 if schedule == 'auto' else 0
        if large_from is not None and not 0 < large_slots < slots:
            raise ValueError('large_slots must leave at least one slot for each class')
        self.large_from = large_from
        nlarge = large_slots if large_from is not None else 0
        self.slots = [_Slot(device, 1 if k < nlarge else 0) for k in range(slots)]
        for slot in self.slots:
            _lib.check(self.lib.ttals_set_schedule(slot.ctx, 1 if schedule == 'auto' else 0,
                                                   1 if gang_precision == 'mixed' else 0), slot.ctx, 'schedule')
            _lib.check(self.lib.ttals_set_time_budget(slot.ctx, float(budget_ms or 0.0)), slot.ctx, 'budget')
            if cost_model is not None:
                p = np.ascontiguousarray(cost_model, dtype=np.float64)
                _lib.check(self.lib.ttals_set_cost_model(slot.ctx, p.ctypes.data, int(p.size)), slot.ctx, 'model')
        self.max_batch = max_batch
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(device)
            hbm_budget = int(0.8 * free)
        self.hbm_budget = hbm_budget
        self.in_use = 0
        self.tol, self.maxit = tol, maxit
        self.queue = collections.deque()
        self.done = []
        self._rid = 0
        self.launches = 0

    # -- requests -------------------------------------------------------------------

    def submit_sle(self, A, x0, b, repeats=1, solver='solve'):
        if solver not in _SLE_SOLVERS:
            raise NotImplementedError("solver must be 'solve' or 'lu'")
        _engine.check_tts([A, x0, b])
        return self._enqueue('sle', ('sle', int(repeats), solver), A, x0, b)

    def submit_evp(self, A, x0, repeats=1, conv_eps=1e-10, solver='eig', sigma=1.0, adaptive_tol=False):
        """Queue an evp.als solve; ``adaptive_tol`` as in ``evp.als_batch``
        (requests are batched only with requests of the same options)."""
        if solver not in _EVP_SOLVERS:
            raise NotImplementedError("solver must be 'eig' or 'eigh'")
        _engine.check_tts([A, x0])
        key = ('evp', int(repeats), float(conv_eps), solver, float(sigma), bool(adaptive_tol))
        return self._enqueue('evp', key, A, x0, None)

    def _enqueue(self, kind, key, A, x0, b):
        nb = footprint_bytes(kind, A, x0, b, self.gang_max)
        if nb > self.hbm_budget:
            raise MemoryError('request needs %d bytes, more than the HBM budget %d' % (nb, self.hbm_budget))
        cls = 0
        if self.large_from is not None:
            r, n = x0.ranks, x0.row_dims
            cls = 1 if max(r[i] * n[i] * r[i + 1] for i in range(x0.order)) >= self.large_from else 0
        t = Ticket(self._rid, kind, key, A, x0, b, nb, cls)
        self._rid += 1
        self.queue.append(t)
        return t

    # -- scheduling -----------------------------------------------------------------

    def poll(self):
        """Retire finished launches, then admit waiting requests into free slots.

        Returns the number of requests completed by this call.
        """
        finished = 0
        for slot in self.slots:
            if slot.busy and _lib.check(self.lib.ttals_query(slot.ctx), slot.ctx, 'query') == 1:
                finished += self._retire(slot)
        for slot in self.slots:
            if not self.queue:
                break
            if not slot.busy:
                self._admit(slot)
        return finished

    def _admit(self, slot):
        head = next((t for t in self.queue if t.cls == slot.cls), None)
        if head is None:
            return
        take = []
        budget = self.hbm_budget - self.in_use
        rest = collections.deque()
        while self.queue and len(take) < self.max_batch:
            t = self.queue.popleft()
            if t.key == head.key and t.cls == head.cls and t.bytes <= budget:
                take.append(t)
                budget -= t.bytes
            else:
                rest.append(t)
                if t.key == head.key and t.cls == head.cls:
                    break  # FIFO within a kind: do not overtake a request that does not fit
        self.queue.extendleft(reversed(rest))
        if not take:
            return
        kind, key = head.kind, head.key
        count = len(take)
        dA, A = _engine.descs([t.A for t in take])
        dX, X = _engine.descs([t.x0 for t in take])
        cores, opacks, O = _engine.outputs([t.x0 for t in take])
        stream = ctypes.c_void_p(slot.stream.cuda_stream)
        # the input tensors were produced on other streams: order the launch after them
        slot.stream.wait_stream(torch.cuda.current_stream(self.device))
        if kind == 'sle':
            dB, B = _engine.descs([t.b for t in take])
            rc = self.lib.ttals_sle_submit(slot.ctx, stream, count, A, X, B, O, key[1], _SLE_SOLVERS[key[2]],
                                           self.tol, self.maxit)
            keep = (dA, dX, dB, A, X, B, O)  # descriptor owners stay alive until ttals_wait
        else:
            policy = evp_mod.ADAPTIVE_CAP_EXP if key[5] else 0
            _lib.check(self.lib.ttals_set_local_tol_policy(slot.ctx, policy), slot.ctx, 'tol policy')
            rc = self.lib.ttals_evp_submit(slot.ctx, stream, count, A, X, O, key[1], key[2], _EVP_SOLVERS[key[3]],
                                           key[4], self.tol, self.maxit)
            keep = (dA, dX, A, X, O)
        _lib.check(rc, slot.ctx, 'submit')
        now = time.perf_counter()
        for t in take:
            t.t_start = now
        self.in_use += sum(t.bytes for t in take)
        slot.batch = (take, cores, opacks, keep)
        self.launches += 1

    def _retire(self, slot):
        take, cores, opacks, _ = slot.batch
        res = (_lib.Result * len(take))()
        rc = self.lib.ttals_wait(slot.ctx, res)
        _lib.check(rc, slot.ctx, 'wait')
        now = time.perf_counter()
        for k, t in enumerate(take):
            info = _engine.SolveInfo(res[k])
            x = _engine.finish(cores, k, opacks, t.x0)
            t.info = info
            t.result = x if t.kind == 'sle' else (info.eigenvalue, x, info.sweeps)
            t.t_done = now
            self.done.append(t)
        self.in_use -= sum(t.bytes for t in take)
        slot.batch = None
        return len(take)

    def busy(self):
        return bool(self.queue) or any(s.busy for s in self.slots)

    def drain(self, timeout=None):
        """Poll until every queued request has completed."""
        t0 = time.perf_counter()
        while self.busy():
            self.poll()
            if timeout is not None and time.perf_counter() - t0 > timeout:
                raise TimeoutError('server did not drain within %.1f s' % timeout)
            time.sleep(2e-4)

    def latency_stats(self):
        lat = np.array([t.latency for t in self.done])
        if lat.size == 0:
            return {}
        return {'completed': int(lat.size), 'p50_ms': float(1e3 * np.percentile(lat, 50)),
                'p99_ms': float(1e3 * np.percentile(lat, 99)), 'launches': self.launches}

    def close(self):
        self.drain()
        for s in self.slots:
            self.lib.ttals_ctx_destroy(s.ctx)
        self.slots = []
