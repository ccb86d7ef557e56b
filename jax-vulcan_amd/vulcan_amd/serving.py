"""Continuous-batching steady-state solver service on one GPU.

Requests are columns (a ``synth.Column``-like object: nz, ni, y, k, atm,
net, dt) to be integrated to steady state with the device step control of
``outer_loop.SteadyStateRunner``.  A request may carry a ``rate_source``
(``RateSource``: its rate network and NASA-9 table) instead of k: its rates
are then built on the device, in its slot, from the T and M of its own
atmosphere (the reference's build_rate_array, rates.py:185-208).  The reference runs one column per
``OuterLoop`` call (outer_loop.py:1112-1240) or a vmapped uniform batch; the
service keeps many columns of different sizes in flight at once:

* one ``SlotPool`` per 16-species bucket (SL = ceil(ni/16)), each a slot
  batch (``vk_batch_slots``) of fixed per-slot capacity on its own HIP
  stream, so the buckets' kernels overlap on the device;
* admission: a request waits in its bucket's FIFO until a slot is free; the
  number of slots per pool is fixed up front from an HBM budget;
* a placed request is uploaded into its slot, the slot descriptor is
  rewritten on the stream (``vk_batch_set_column``) and only that slot is
  initialised (``vk_ss_init_range_f64``); the other slots keep running;
* every pool advances ``chunk`` attempts at a time (optionally one HIP
  graph replay); finished slots (status CONVERGED or MAXSTEPS) are read
  back and refilled.

All per-column arithmetic is independent of the other slots and of the slot
a column lands in, so a request's result is the same alone and in any mix.
"""
from __future__ import annotations

import collections
import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Dict, List, NamedTuple, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .core import F64, DeviceNetwork, atm_record_layout, block_doubles, device, pack_atm_record
from .outer_loop import (ST_CONVERGED, ST_DONE, ST_MAXSTEPS, VK_MAX_ATOMS, RunConfig, SsState)

FACTOR_CR = 6      # vk_tune_set(VK_TUNE_FACTOR, sl, VK_FACTOR_CR)

class RateSource(NamedTuple):
    """What a request needs to have its k built on the device."""
    net: object                      # rate network (RateNetwork or network.Network)
    nasa9: np.ndarray                # (ni, 2, 10)
    use_lowT_limit_rates: bool = False
    remove_list: Optional[Sequence[int]] = None


ST_IDLE = 16
IDLE_STATUS = ST_IDLE | ST_CONVERGED  # DONE bits set: every kernel skips the slot


@dataclass
class Result:
    """Outcome of one request (times in seconds from the start of the run)."""
    index: int
    arrival: float
    start: float
    finish: float
    status: int
    attempts: int
    accept: int
    y: np.ndarray
    ymix: np.ndarray

    @property
    def latency(self) -> float:
        return self.finish - self.arrival

    @property
    def converged(self) -> bool:
        return bool(self.status & ST_CONVERGED)


def _bucket(ni: int) -> int:
    return (int(ni) + 15) // 16


class SlotPool:
    """Fixed-capacity slots for one 16-species bucket, on its own stream."""

    def __init__(self, nslots: int, cap_nz: int, cap_ni: int, cap_S: int, cap_T: int,
                 cfg: RunConfig, use_graph: bool = True, chunk: int = 20,
                 latency_mode: bool = True):
        lib = _lib.load()
        self.dev = device()
        self.nslots, self.cap_nz, self.cap_ni, self.cap_S, self.cap_T = (
            int(nslots), int(cap_nz), int(cap_ni), int(cap_S), int(cap_T))
        h = C.c_void_p()
        _lib.check(lib.vk_batch_slots(self.nslots, self.cap_nz, self.cap_ni, self.cap_S,
                                      self.cap_T, C.byref(h)), "vk_batch_slots")
        self.handle = h
        # the cyclic-reduction factor for this pool only (before the
        # workspace is sized)
        _lib.check(lib.vk_batch_set_latency_mode(h, int(bool(latency_mode))),
                   "vk_batch_set_latency_mode")
        self.stream = torch.cuda.Stream(device=self.dev)  # default priority
        n, dev = self.nslots, self.dev
        self.cs = self.cap_nz * self.cap_ni
        self.ck = self.cap_S * self.cap_nz
        self.ca = atm_record_layout(self.cap_nz, self.cap_ni)[0]
        with torch.cuda.stream(self.stream):
            z = lambda m: torch.zeros(m, dtype=F64, device=dev)
            self.y, self.y_prev, self.ymix, self.sol, self.delta_arr = (z(n * self.cs) for _ in range(5))
            self.compo = torch.ones(n * self.cs, dtype=F64, device=dev)  # one "atom" per species
            self.k = z(n * self.ck)
            self.atm = z(n * self.ca)
            self.atom_ini, self.atom_loss, self.atom_loss_prev = (z(n * VK_MAX_ATOMS) for _ in range(3))
            self.dt, self.t, self.delta, self.rel = (z(n) for _ in range(4))
            zi = lambda: torch.zeros(n, dtype=torch.int32, device=dev)
            (self.accept, self.retry, self.nega, self.loss, self.dcount, self.attempts,
             self.photo_due) = (zi() for _ in range(7))
            self.status = torch.full((n,), IDLE_STATUS, dtype=torch.int32, device=dev)
            nbytes = int(lib.vk_ros2_workspace_bytes(self.handle))
            self.ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        self.params = cfg.c_params(1, False)
        self.state = SsState(*[t.data_ptr() for t in (
            self.y, self.y_prev, self.ymix, self.sol, self.delta_arr, self.compo, self.atom_ini,
            self.atom_loss, self.atom_loss_prev, self.dt, self.t, self.delta, self.rel,
            self.accept, self.retry, self.nega, self.loss, self.dcount, self.status,
            self.attempts, self.photo_due)], None)
        self.free: List[int] = list(range(n))[::-1]
        self.busy: Dict[int, tuple] = {}       # slot -> (request index, column, start time)
        self.nets: Dict[int, DeviceNetwork] = {}
        self.retired: List[DeviceNetwork] = []  # released after the stream's next sync
        self.chunk = int(chunk)
        self.use_graph = bool(use_graph)
        self.graph = None
        self.queue: collections.deque = collections.deque()
        self._fin_ptrs = lib  # keep the library alive for the finalizer
        import weakref
        self._fin = weakref.finalize(self, lib.vk_batch_destroy, h)

    @staticmethod
    def bytes_per_slot(cap_nz: int, cap_ni: int, cap_S: int, latency_mode: bool = True) -> int:
        cs, ck = cap_nz * cap_ni, cap_S * cap_nz
        ca = atm_record_layout(cap_nz, cap_ni)[0]
        # Ros2 workspace per slot: the diagonal blocks, their inverses, vectors
        ws = 2 * cap_nz * block_doubles(cap_ni) + 16 * cs + ck
        sl = (cap_ni + 15) // 16
        if latency_mode or _lib.load().vk_tune_get(0, sl) == FACTOR_CR:
            ws += 7 * cap_nz * block_doubles(cap_ni)      # the cyclic-reduction levels
        return 8 * (6 * cs + ck + ca + ws) + 8 * (3 * VK_MAX_ATOMS + 4) + 4 * 9

    # ------------------------------------------------------------------
    def place(self, slot: int, col, index: int, now: float):
        lib = _lib.load()
        nz, ni = int(col.nz), int(col.ni)
        dn = DeviceNetwork(col.net)
        st = self.stream.cuda_stream
        with torch.cuda.stream(self.stream):
            o = slot * self.cs
            self.y[o:o + nz * ni].copy_(torch.as_tensor(np.ascontiguousarray(col.y, np.float64)).reshape(-1))
            ok = slot * self.ck
            oa = slot * self.ca
            rec = pack_atm_record(col.atm)
            self.atm[oa:oa + rec.numel()].copy_(rec)
            src = getattr(col, "rate_source", None)
            if src is not None:
                from . import rates
                _, off = atm_record_layout(nz, ni)
                d = rates.device_rates(src.net, src.nasa9,
                                       rates.remove_mask(src.net, src.remove_list, paired=False))
                d.eval(rates.RATES_BUILD, self.atm[oa + off[5]:oa + off[5] + nz],
                       self.atm[oa + off[7]:oa + off[7] + nz],
                       lowt_on=src.use_lowT_limit_rates,
                       out=self.k[ok:ok + dn.S * nz].view(dn.S, nz))
            else:
                self.k[ok:ok + dn.S * nz].copy_(
                    torch.as_tensor(np.ascontiguousarray(col.k, np.float64)).reshape(-1))
            self.dt[slot:slot + 1].fill_(float(col.dt))
            _lib.check(lib.vk_batch_set_column(self.handle, slot, nz, dn.handle, st),
                       "vk_batch_set_column")
            _lib.check(lib.vk_ss_init_range_f64(self.handle, slot, 1, self.atm.data_ptr(),
                                                C.byref(self.params), C.byref(self.state), st),
                       "vk_ss_init_range_f64")
        self.nets[slot] = dn
        self.busy[slot] = (index, col, now)

    def _attempt(self):
        lib = _lib.load()
        _lib.check(lib.vk_ss_attempt_f64(self.handle, self.k.data_ptr(), self.atm.data_ptr(),
                                         C.byref(self.params), C.byref(self.state),
                                         self.ws.data_ptr(), self.ws.numel(),
                                         self.stream.cuda_stream), "vk_ss_attempt_f64")

    def advance(self):
        """Queue `chunk` attempts on the pool's stream (asynchronous)."""
        if self.use_graph:
            if self.graph is None:
                with torch.cuda.stream(self.stream):
                    self._attempt()  # first launch outside capture (one-time setup)
                self.stream.synchronize()
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=self.stream):
                    for _ in range(self.chunk):
                        self._attempt()
                self.graph = g
                return
            with torch.cuda.stream(self.stream):
                self.graph.replay()
        else:
            with torch.cuda.stream(self.stream):
                for _ in range(self.chunk):
                    self._attempt()

    def finished(self) -> List[int]:
        """Busy slots whose status says done (synchronises the pool's stream)."""
        with torch.cuda.stream(self.stream):
            st = self.status.cpu().numpy()
        self.retired.clear()
        return [s for s in self.busy if st[s] & ST_DONE]

    def harvest(self, slot: int, now: float) -> Result:
        index, col, start = self.busy.pop(slot)
        nz, ni = int(col.nz), int(col.ni)
        o = slot * self.cs
        with torch.cuda.stream(self.stream):
            y = self.y[o:o + nz * ni].reshape(nz, ni).cpu().numpy()
            ymix = self.ymix[o:o + nz * ni].reshape(nz, ni).cpu().numpy()
            status = int(self.status[slot].item())
            attempts = int(self.attempts[slot].item())
            accept = int(self.accept[slot].item())
            self.status[slot:slot + 1].fill_(IDLE_STATUS)
        self.retired.append(self.nets.pop(slot))
        self.free.append(slot)
        return Result(index, col._arrival, start, now, status, attempts, accept, y, ymix)


class ServingEngine:
    """Continuous batching of steady-state solves within an HBM budget.

    ``capacity`` maps a 16-species bucket SL to (cap_nz, cap_ni, cap_S,
    cap_T); by default it is derived from the requests given to ``run``.
    Each bucket gets ``min(max_slots, budget share / bytes per slot)``
    slots, the budget shared in proportion to the bucket's request count.
    Requests stop by the relative-change rule (RunConfig.conv_mode
    "relmax"); the ring of the longdy rule is not kept per slot.
    ``latency_mode`` (default on) gives every pool the cyclic-reduction
    block factor (vk_batch_set_latency_mode): a pool holds a few columns per
    bucket, where one workgroup per column would leave the GPU idle.
    """

    def __init__(self, hbm_budget_bytes: float = 200e9, max_slots: int = 512,
                 cfg: RunConfig = RunConfig(), chunk: int = 20, use_graph: bool = True,
                 capacity: Optional[Dict[int, tuple]] = None, latency_mode: bool = True):
        if cfg.conv_mode != "relmax":
            raise ValueError("serving pools stop requests by the relative-change rule "
                             "(conv_mode 'relmax')")
        self.latency_mode = bool(latency_mode)
        self.budget = float(hbm_budget_bytes)
        self.max_slots = int(max_slots)
        self.cfg = cfg
        self.chunk = int(chunk)
        self.use_graph = use_graph
        self.capacity = capacity
        self.pools: Dict[int, SlotPool] = {}
        self.slot_counts: Dict[int, int] = {}

    def _build_pools(self, requests: Sequence):
        caps = dict(self.capacity or {})
        counts = collections.Counter(_bucket(r.ni) for r in requests)
        if not self.capacity:
            for r in requests:
                sl = _bucket(r.ni)
                S, T = int(r.net.nr) + 1, int(np.asarray(r.net.reactant_idx).shape[1])
                c = caps.get(sl, (0, 0, 0, 0))
                caps[sl] = (max(c[0], int(r.nz)), max(c[1], int(r.ni)), max(c[2], S), max(c[3], T))
        total = sum(counts.values()) or 1
        for sl, (cnz, cni, cS, cT) in sorted(caps.items()):
            share = self.budget * counts.get(sl, 0) / total if counts else self.budget / len(caps)
            per = SlotPool.bytes_per_slot(cnz, cni, cS, self.latency_mode)
            n = int(min(self.max_slots, max(1, share // per), max(1, counts.get(sl, 1))))
            self.pools[sl] = SlotPool(n, cnz, cni, cS, cT, self.cfg, self.use_graph, self.chunk,
                                      self.latency_mode)
            self.slot_counts[sl] = n

    def run(self, requests: Sequence, arrivals: Optional[Sequence[float]] = None,
            realtime: bool = True, max_rounds: int = 10 ** 9,
            report_every: float = 0.0, metrics=None) -> List[Result]:
        """Serve the requests; arrival times in seconds from the start (all 0 if None).

        With realtime=True a request is admitted only once the wall clock has
        reached its arrival time; latency = finish - arrival.  report_every > 0
        prints a progress line (completed / in flight / queued) after the
        chunk that crosses each such interval.  ``metrics`` (a path, stream or
        metrics.JsonlLog) receives a ``request_done`` record per finished
        request and a ``serving_summary`` record at the end."""
        from .metrics import open_log
        log = open_log(metrics)
        reqs = list(requests)
        arr = np.zeros(len(reqs)) if arrivals is None else np.asarray(arrivals, dtype=np.float64)
        for i, r in enumerate(reqs):
            r._arrival = float(arr[i])
        if not self.pools:
            self._build_pools(reqs)
        order = np.argsort(arr, kind="stable")
        pending = collections.deque(int(i) for i in order)
        results: List[Optional[Result]] = [None] * len(reqs)
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float("inf"))
        rounds = 0
        next_report = report_every
        while (pending or any(p.busy or p.queue for p in self.pools.values())) and rounds < max_rounds:
            rounds += 1
            now = clock()
            while pending and arr[pending[0]] <= now:
                i = pending.popleft()
                self.pools[_bucket(reqs[i].ni)].queue.append(i)
            t_now = time.perf_counter() - t0
            for p in self.pools.values():
                while p.queue and p.free:
                    i = p.queue.popleft()
                    p.place(p.free.pop(), reqs[i], i, t_now)
            active = [p for p in self.pools.values() if p.busy]
            if not active:
                if pending:
                    time.sleep(max(0.0, min(0.05, arr[pending[0]] - clock())))
                continue
            for p in active:
                p.advance()
            for p in active:
                done = p.finished()
                t_done = time.perf_counter() - t0
                for s in done:
                    res = p.harvest(s, t_done)
                    results[res.index] = res
                    if log is not None:
                        log.write("request_done", index=res.index, bucket=_bucket(reqs[res.index].ni),
                                  latency_s=round(res.latency, 6),
                                  wait_s=round(res.start - res.arrival, 6),
                                  attempts=res.attempts, accepted=res.accept,
                                  status=res.status, converged=res.converged)
            if report_every > 0 and time.perf_counter() - t0 >= next_report:
                next_report += report_every
                n_done = sum(r is not None for r in results)
                n_busy = sum(len(p.busy) for p in self.pools.values())
                n_q = sum(len(p.queue) for p in self.pools.values()) + len(pending)
                print(f"[serve] t={time.perf_counter() - t0:.0f}s done={n_done} in_flight={n_busy} "
                      f"waiting={n_q} chunks={rounds}", flush=True)
        for p in self.pools.values():
            p.stream.synchronize()
        out = [r for r in results if r is not None]
        if log is not None:
            log.write("serving_summary", slots=dict(self.slot_counts), **latency_summary(out))
        return out


def latency_summary(results: Sequence[Result]) -> dict:
    lat = np.array([r.latency for r in results])
    if lat.size == 0:
        return {}
    span = max(r.finish for r in results) - min(r.arrival for r in results)
    return {"n": int(lat.size), "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
            "throughput_solves_per_s": float(lat.size / span) if span > 0 else None,
            "converged": int(sum(r.converged for r in results)),
            "attempts_mean": float(np.mean([r.attempts for r in results]))}
