"""Continuous-batching reconstruction serving on one GPU.

Requests of different sizes and geometries are admitted through a device-
memory ledger; each admitted request keeps its volume, error sinogram,
Hessian and VCD scratch resident on the GPU and becomes a job of the native
request engine (mbir_server_* in libmbir_hip.so), whose worker thread
advances every active job by one subset step per round on the job's own
stream, so the steps of different reconstructions overlap on the device.
Requests that arrive while others run are admitted as memory and slots free
up.  Each job runs exactly the subset steps of a solo ``recon`` with the same
seed, so a result is the same alone and in the mix.

The reference has no request loop; the closed-form ledger follows the
structure of its per-phase memory accounting (_memory_ledger.py:241-375).
"""

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from . import vcd_utils
from .parallel_beam import ParallelBeamModel


class MemoryPreflightError(RuntimeError):
    """A request does not fit the device-memory budget even when alone."""


def request_device_bytes(sino_shape, recon_shape, granularity, weighted=False):
    """Device bytes of one request by phase (closed form).

    resident: error sinogram, recon, Hessian diagonal, (weights), partition
    index tables, per-iteration statistics.  init: the transient peak while
    the request is prepared (measured sinogram + filtered sinogram + Hessian
    weights + the full-region cylinders).  step: VCD scratch of the largest
    subset (back-projection partials for up to one split per 8 views,
    direction, partials) plus the stored A(delta) (one float32 sinogram).
    """
    V, R, C = (int(v) for v in sino_shape)
    Nr, Nc, S = (int(v) for v in recon_shape)
    sino = 4 * V * R * C
    vol = 4 * Nr * Nc * S
    P = int(vcd_utils.get_2d_ror_mask(recon_shape).sum())
    parts = 0
    max_p = 0
    for g in granularity:
        n = max(1, min(int(g), P))
        pk = -(-P // n)
        parts += 8 * n * pk
        max_p = max(max_p, pk)
    resident = sino + 2 * vol + (sino if weighted else 0) + parts + 8 * 64 * (S + 6)
    init = 3 * sino + 4 * P * S
    splits = max(1, V // 8)
    step = (4 * max_p * S * (splits + 2) + sino + 8 * 2 * (V * C + 4 * max_p * S // 256 + 4096))
    return {'resident': resident, 'init': init, 'step': step,
            'total': resident + max(init, step)}


class MemoryLedger:
    """Admission by device bytes: a request is admitted when its total fits
    the budget minus a safety margin and the bytes already held."""

    def __init__(self, budget_bytes, margin=0.15):
        self.capacity = int(budget_bytes * (1.0 - margin))
        self.used = 0

    def fits_alone(self, nbytes):
        return nbytes <= self.capacity

    def try_reserve(self, nbytes):
        if self.used + nbytes > self.capacity:
            return False
        self.used += nbytes
        return True

    def release(self, nbytes):
        self.used -= nbytes
        assert self.used >= 0


@dataclass
class ReconRequest:
    """One reconstruction: a parallel-beam sinogram (views, rows, channels)
    and its angles, recon parameters for set_params, the iteration count and
    the numpy seed a solo ``np.random.seed(seed); recon(...)`` would use."""
    sinogram: np.ndarray
    angles: np.ndarray
    max_iterations: int = 20
    seed: int = 0
    params: dict = field(default_factory=dict)
    tag: int = 0
    arrival: float = 0.0


@dataclass
class ReconResult:
    tag: int
    recon: np.ndarray
    fm_rmse: np.ndarray
    nmae: np.ndarray
    alpha: np.ndarray
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self):
        return self.finished - self.arrival


def _summaries(stats, S, total_sino, weighted, sigma_y, num_subsets):
    """Per-iteration fm_rmse, nmae and mean alpha from the stats blocks."""
    fm, nmae, alpha = [], [], []
    for h, n_sub in zip(stats, num_subsets):
        wsq, wsum = h[S + 2], h[S + 4]
        avg_w = wsum / total_sino if weighted else 1.0
        fm.append(math.sqrt(wsq / (avg_w * total_sino)) / sigma_y)
        nmae.append(h[S] / h[S + 5] if h[S + 5] else float('nan'))
        alpha.append(h[S + 1] / n_sub)
    return np.array(fm), np.array(nmae), np.array(alpha)


class ReconServer:
    """The request loop: submit(), poll(), close().  ``max_active`` jobs run
    at once; ``budget_bytes`` (default: 85% of the device's free memory at
    construction) bounds what is admitted."""

    def __init__(self, max_active=8, budget_bytes=None, device=None):
        if not torch.cuda.is_available():
            raise RuntimeError('ReconServer needs a HIP device (there is no CPU path)')
        dev = torch.device('cuda' if device is None else device)
        if dev.index is None:
            dev = torch.device('cuda', torch.cuda.current_device())
        self.device = dev
        torch.cuda.set_device(self.device)
        if budget_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            budget_bytes = int(0.85 * free)
        self.ledger = MemoryLedger(budget_bytes)
        self.lib = _lib.lib()
        self.handle = self.lib.mbir_server_create(int(max_active))
        if not self.handle:
            raise RuntimeError('mbir_server_create failed')
        self.queue = []
        self.jobs = {}
        self._tags = (ctypes.c_int64 * 64)()
        self._status = (ctypes.c_int32 * 64)()

    # -- admission ---------------------------------------------------------
    def submit(self, req):
        """Queue a request (admitted as soon as the ledger allows)."""
        model = ParallelBeamModel(req.sinogram.shape, req.angles)
        if req.params:
            model.set_params(no_warning=True, **req.params)
        recon_shape, gran = model.get_params(['recon_shape', 'granularity'])
        need = request_device_bytes(req.sinogram.shape, recon_shape, gran)['total']
        if not self.ledger.fits_alone(need):
            raise MemoryPreflightError(
                f'request {req.tag} needs {need / 2**30:.2f} GiB; the budget is '
                f'{self.ledger.capacity / 2**30:.2f} GiB')
        self.queue.append((req, model, need))
        self._admit()

    def _admit(self):
        while self.queue and self.ledger.try_reserve(self.queue[0][2]):
            req, model, need = self.queue.pop(0)
            self._start(req, model, need)

    def _start(self, req, model, need):
        admitted = time.perf_counter()
        saved = np.random.get_state()
        try:
            np.random.seed(req.seed)
            (sino, weights, init, partitions, seq, _gran,
             _reg) = model.initialize_recon(req.sinogram, max_iterations=req.max_iterations,
                                            logfile_path=None, print_logs=False)
            with torch.no_grad():
                engine, _hess = model._prepare_vcd(sino, partitions, weights=weights,
                                                   init_recon=init)
            # The same draws, in the same order, as the solo run's iterations.
            orders = [np.random.permutation(int(partitions[int(p)].shape[0])) for p in seq]
        finally:
            np.random.set_state(saved)
        schedule = np.concatenate(
            [o.astype(np.int64) + int(engine.offsets[int(p)]) for o, p in zip(orders, seq)]
            + [np.zeros(0, np.int64)]).astype(np.int32)
        offsets = np.zeros(len(seq) + 1, np.int32)
        offsets[1:] = np.cumsum([len(o) for o in orders])
        S = engine.S
        stats = torch.zeros((len(seq), S + 6), dtype=torch.float64, device=self.device)
        # The job's stream must see the preparation done on this stream.
        torch.cuda.current_stream(self.device).synchronize()
        st = self.lib.mbir_server_submit(
            self.handle, ctypes.byref(engine.job), schedule.ctypes.data_as(ctypes.c_void_p),
            offsets.ctypes.data_as(ctypes.c_void_p), len(seq), _lib.ptr(stats), int(req.tag))
        _lib.check(st, 'server submit')
        self.jobs[int(req.tag)] = dict(req=req, need=need, model=model, engine=engine,
                                       schedule=schedule, offsets=offsets, stats=stats,
                                       seq=seq, partitions=partitions, admitted=admitted,
                                       weighted=weights is not None)

    # -- completion ----------------------------------------------------------
    def poll(self, timeout_ms=0):
        """Finished requests as ReconResult (admitting queued ones)."""
        n = self.lib.mbir_server_poll(self.handle, self._tags, self._status, 64,
                                      int(timeout_ms))
        if n < 0:
            _lib.check(n, 'server poll')
        out = []
        failed = []
        for i in range(n):
            tag, status = int(self._tags[i]), int(self._status[i])
            job = self.jobs.pop(tag)
            self.ledger.release(job['need'])
            if status != 0:
                failed.append((tag, status))
                continue
            out.append(self._finish(job))
        self._admit()
        for tag, status in failed:  # after the ledger and the queue are settled
            _lib.check(status, f'request {tag}')
        return out

    def _finish(self, job):
        model, engine, req = job['model'], job['engine'], job['req']
        rs = model.get_params('recon_shape')
        recon = engine.x.reshape(rs[0], rs[1], rs[2]).cpu().numpy()
        h = job['stats'].cpu().numpy()
        n_sub = [int(job['partitions'][int(p)].shape[0]) for p in job['seq']]
        fm, nmae, alpha = _summaries(h, engine.S, math.prod(req.sinogram.shape),
                                     job['weighted'], model.get_params('sigma_y'), n_sub)
        return ReconResult(tag=int(req.tag), recon=recon, fm_rmse=fm, nmae=nmae, alpha=alpha,
                           arrival=req.arrival, admitted=job['admitted'],
                           finished=time.perf_counter())

    def pending(self):
        return len(self.queue) + len(self.jobs)

    def close(self):
        if self.handle:
            self.lib.mbir_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- a request stream ------------------------------------------------------
    def run_stream(self, requests, time_scale=1.0):
        """Serve requests whose ``arrival`` fields are offsets in seconds from
        now (scaled by time_scale); returns results in completion order, with
        arrival/finish times on the perf_counter clock."""
        t0 = time.perf_counter()
        todo = sorted(requests, key=lambda r: r.arrival)
        for r in todo:
            r.arrival = t0 + r.arrival * time_scale
        results = []
        i = 0
        while i < len(todo) or self.pending():
            now = time.perf_counter()
            while i < len(todo) and todo[i].arrival <= now:
                self.submit(todo[i])
                i += 1
            wait = 0 if i < len(todo) else 5
            results.extend(self.poll(timeout_ms=wait))
            if i < len(todo) and not results and self.pending() == 0:
                time.sleep(max(0.0, min(0.002, todo[i].arrival - time.perf_counter())))
        return results


def run_plain(requests, time_scale=1.0):
    """The one-at-a-time path: each request, in arrival order, through
    ParallelBeamModel.recon (same seed); results carry the same timings."""
    t0 = time.perf_counter()
    todo = sorted(requests, key=lambda r: r.arrival)
    results = []
    for r in todo:
        arrival = t0 + r.arrival * time_scale
        now = time.perf_counter()
        if now < arrival:
            time.sleep(arrival - now)
        start = time.perf_counter()
        model = ParallelBeamModel(r.sinogram.shape, r.angles)
        if r.params:
            model.set_params(no_warning=True, **r.params)
        np.random.seed(r.seed)
        recon, d = model.recon(r.sinogram, max_iterations=r.max_iterations,
                               stop_threshold_change_pct=0.0, logfile_path=None,
                               print_logs=False)
        rp = d['recon_params']
        results.append(ReconResult(tag=int(r.tag), recon=recon,
                                   fm_rmse=np.array(rp['fm_rmse']),
                                   nmae=np.array(rp['stop_threshold_change_pct']) / 100.0,
                                   alpha=np.array(rp['alpha_values']), arrival=arrival,
                                   admitted=start, finished=time.perf_counter()))
    return results


def latency_summary(results):
    """recons/s over the makespan and p50/p99 latency (seconds)."""
    lat = np.array([r.latency for r in results])
    t_first = min(r.arrival for r in results)
    t_last = max(r.finished for r in results)
    return {'n': len(results), 'recons_per_s': len(results) / (t_last - t_first),
            'p50_s': float(np.percentile(lat, 50)), 'p99_s': float(np.percentile(lat, 99)),
            'makespan_s': t_last - t_first}
