"""Adaptive steady-state runs with the step control on the device.

``SteadyStateRunner`` is the GPU counterpart of the reference's
``OuterLoop`` runner for the step-control part of ``body_fn`` /
``cond_fn`` (outer_loop.py:492-1112): every column of a batch keeps its
state, step size, counters and status on the device, one *attempt* is a
Ros2 step plus the controller for every unfinished column, and attempts
can be replayed from a captured HIP graph.  The host only polls the status
words every ``check_every`` attempts.

Photolysis, when the columns carry a ``PhotoSetup``, refreshes the
photolysis rows of k before an attempt with retry_count == 0 and
accept_count % photo_frq == 0 (outer_loop.py:716-724).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import NamedTuple, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .core import F64, Batch, device, device_network, pack_atm_record

VK_MAX_ATOMS = 8
ST_SINGULAR, ST_NONFINITE, ST_CONVERGED, ST_MAXSTEPS = 1, 2, 4, 8
ST_DONE = ST_CONVERGED | ST_MAXSTEPS


class SsParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "rtol", "dt_var_min", "dt_var_max", "dt_min", "dt_max", "atol", "mtol", "pos_cut",
        "nega_cut", "loss_eps", "trun_min", "conv_tol", "flux_cri")] + [(n, C.c_int32) for n in (
            "batch_max_retries", "count_min", "count_max", "na", "zero_bot_row", "photo_frq",
            "update_frq", "conv_mode", "conv_step", "conv_stall_window")] + [
                (n, C.c_double) for n in ("st_factor", "yconv_cri", "slope_cri", "yconv_min",
                                          "mtol_conv", "runtime")] + [
                    ("photo_frq_final", C.c_int32), ("photo_sw_longdy", C.c_double),
                    ("photo_sw_longdydt", C.c_double)]


class SsState(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "y", "y_prev", "ymix", "sol", "delta_arr", "compo", "atom_ini", "atom_loss",
        "atom_loss_prev", "dt", "t", "delta", "rel", "accept", "retry", "nega", "loss",
        "dcount", "status", "attempts", "photo_due", "aflux_change", "refresh_due", "pin",
        "off_pin", "ring_y", "ring_t", "conv", "hp", "off_hp", "conv_ign", "off_ign",
        "photo_cad")]


PIN_CHARGE, PIN_ALL_BOT, PIN_SP_BOT, PIN_H2HE = 1, 2, 4, 8


class ColumnConstraints(NamedTuple):
    """Constraints on every accepted state of one column, applied after the
    hydrostatic balance (outer_loop.py:855-869): ion charge balance
    (charge (ni) with 0 at e_idx, e = -sum_i charge_i y_i), the whole bottom
    layer fixed (bottom_n (ni) densities), and pinned bottom mixing ratios
    (fix_sp_bot {species index: mixing ratio}, times the bottom density), and
    the Hycean H2/He pin (h2he = (h2_idx, he_idx, pin_time): once t exceeds
    pin_time the bottom H2 and He keep that moment's mixing ratios,
    outer_loop.py:871-895)."""
    charge: Optional[np.ndarray] = None
    e_idx: int = -1
    bottom_n: Optional[np.ndarray] = None
    fix_sp_bot: Optional[dict] = None
    h2he: Optional[tuple] = None

    def record(self, ni: int) -> np.ndarray:
        flags = 0
        rec = np.zeros(2 + 3 * ni + 6)
        rec[2 + 2 * ni:2 + 3 * ni] = np.nan
        if self.h2he is not None:
            h2, he, t_pin = self.h2he
            if not (0 <= int(h2) < ni and 0 <= int(he) < ni):
                raise ValueError("h2he species outside 0..ni-1")
            rec[2 + 3 * ni:2 + 3 * ni + 3] = (int(h2), int(he), float(t_pin))
            flags |= PIN_H2HE
        if self.charge is not None:
            if not 0 <= int(self.e_idx) < ni:
                raise ValueError(f"e_idx {self.e_idx} outside 0..{ni - 1}")
            q = np.asarray(self.charge, np.float64).reshape(-1)
            if q.shape[0] != ni:
                raise ValueError("charge needs one entry per species")
            rec[2:2 + ni] = q
            rec[2 + int(self.e_idx)] = 0.0
            rec[1] = int(self.e_idx)
            flags |= PIN_CHARGE
        if self.bottom_n is not None:
            rec[2 + ni:2 + 2 * ni] = np.asarray(self.bottom_n, np.float64).reshape(-1)
            flags |= PIN_ALL_BOT
        if self.fix_sp_bot:
            for i, m in self.fix_sp_bot.items():
                if not 0 <= int(i) < ni:
                    raise ValueError(f"pinned species {i} outside 0..{ni - 1}")
                rec[2 + 2 * ni + int(i)] = float(m)
            flags |= PIN_SP_BOT
        rec[0] = flags
        return rec


class ColumnConv(NamedTuple):
    """What the reference's stopping rule reads of one column besides its
    state: the layer scale heights Hp (nz; slope_min, outer_loop.py:676-683)
    and the species longdy ignores (conver_ignore, indices)."""
    Hp: np.ndarray
    ignore: Sequence[int] = ()


@dataclass
class RunConfig:
    """Solver tolerances (vulcan_cfg.py:125-152) and the stopping rule.

    conv_mode "relmax": a column is done when the largest masked relative
    change of an accepted step falls below conv_tol.  "longdy": the
    reference's rule (outer_loop.py:636-713, 918-1003): longdy against a
    ring of the last conv_step accepted states at the entry nearest
    t * st_factor, the strict (yconv_cri, slope_cri) and loose (yconv_min,
    slope_min) criteria, the stall fallback after conv_stall_window steps,
    and the runtime / count_max caps (count_max as accept_count > count_max).
    """
    rtol: float = 0.2
    dt_var_min: float = 0.5
    dt_var_max: float = 2.0
    dt_min: float = 1e-14
    dt_max: float = 1e22 * 1e-5
    atol: float = 0.1
    mtol: float = 1e-22
    pos_cut: float = 0.0
    nega_cut: float = -1.0
    loss_eps: float = 0.1
    batch_max_retries: int = 8
    trun_min: float = 1e2
    count_min: int = 120
    count_max: int = 5000
    conv_tol: float = 1e-4
    flux_cri: float = 0.1
    photo_frq: int = 5           # the reference's ini_update_photo_frq (its default: 100)
    photo_frq_final: int = 0     # final_update_photo_frq (default 5); 0: no switch
    photo_switch_longdy: float = -1.0    # < 0: yconv_min * 10 (outer_loop.py:1327-1330)
    photo_switch_longdydt: float = 1e-6
    update_frq: int = 0          # atmosphere refresh cadence (vulcan_cfg.update_frq); 0: off
    conv_mode: str = "relmax"
    conv_step: int = 500         # vulcan_cfg.py:120-122, 131-147
    st_factor: float = 0.5
    conv_stall_window: int = 200
    yconv_cri: float = 0.01
    slope_cri: float = 1e-4
    yconv_min: float = 0.1
    mtol_conv: float = 1e-20
    runtime: float = 1e22

    def c_params(self, na: int, zero_bot_row: bool = False) -> SsParams:
        p = SsParams()
        for f in ("rtol", "dt_var_min", "dt_var_max", "dt_min", "dt_max", "atol", "mtol",
                  "pos_cut", "nega_cut", "loss_eps", "trun_min", "conv_tol", "flux_cri"):
            setattr(p, f, float(getattr(self, f)))
        p.batch_max_retries = int(self.batch_max_retries)
        p.count_min = int(self.count_min)
        p.count_max = int(self.count_max)
        p.na = int(na)
        p.zero_bot_row = int(bool(zero_bot_row))
        p.photo_frq = int(self.photo_frq)
        p.update_frq = int(self.update_frq)
        if self.conv_mode not in ("relmax", "longdy"):
            raise ValueError(f"conv_mode must be 'relmax' or 'longdy', not {self.conv_mode!r}")
        p.conv_mode = 1 if self.conv_mode == "longdy" else 0
        p.conv_step = int(self.conv_step)
        p.conv_stall_window = int(self.conv_stall_window)
        for f in ("st_factor", "yconv_cri", "slope_cri", "yconv_min", "mtol_conv", "runtime"):
            setattr(p, f, float(getattr(self, f)))
        p.photo_frq_final = int(self.photo_frq_final)
        p.photo_sw_longdy = float(self.photo_switch_longdy if self.photo_switch_longdy >= 0.0
                                  else 10.0 * self.yconv_min)
        p.photo_sw_longdydt = float(self.photo_switch_longdydt)
        return p


class SteadyStateRunner:
    """Batch of adaptive Ros2 runs, uneven nz / ni / networks allowed."""

    def __init__(self, columns: Sequence, cfg: RunConfig = RunConfig(), compo=None,
                 photo: Optional[object] = None, refresh: Optional[Sequence] = None,
                 constraints: Optional[Sequence] = None, conv: Optional[Sequence] = None,
                 latency_mode: bool = False):
        """refresh: one ``atm_refresh.AtmRefreshStatic`` per column; with
        cfg.update_frq > 0 each column's geometry (g, dzi, Hpi) and escape
        fluxes are refreshed in the loop after every update_frq-th accepted
        step (outer_loop.py:809-843).  constraints: one ``ColumnConstraints``
        (or None) per column.  conv: one ``ColumnConv`` per column, needed
        by cfg.conv_mode "longdy" (with the refresh on, slope_min follows
        the refreshed Hp).  latency_mode: the cyclic-reduction block factor
        for this batch (vk_batch_set_latency_mode), for a few columns per
        bucket."""
        self.dev = device()
        self.cfg = cfg
        cols = list(columns)
        self.columns = cols
        nets, net_ids, seen = [], [], {}
        for c in cols:
            key = id(c.net)
            if key not in seen:
                seen[key] = len(nets)
                nets.append(device_network(c.net))
            net_ids.append(seen[key])
        self.batch = Batch([c.nz for c in cols], nets, net_ids)
        b = self.batch
        self.na = 1 if compo is None else int(np.asarray(compo[0]).shape[1])
        if not 1 <= self.na <= VK_MAX_ATOMS:
            raise ValueError(f"na must be in 1..{VK_MAX_ATOMS}")
        dev, n = self.dev, b.ncols
        z = lambda: torch.zeros(b.state_len, dtype=F64, device=dev)
        self.y = b.pack_state([c.y for c in cols], dev)
        self.y_prev, self.ymix, self.sol, self.delta_arr = z(), z(), z(), z()
        comp = torch.zeros(b.state_len, dtype=F64)
        for ci, c in enumerate(cols):
            cm = np.ones((c.ni, 1)) if compo is None else np.asarray(compo[ci], dtype=np.float64)
            o = int(b.off_s[ci])
            comp[o:o + cm.size] = torch.as_tensor(cm.reshape(-1))
        self.compo = comp.to(dev)
        self.k = b.pack_k([c.k for c in cols], dev)
        self.atm = b.pack_atm([c.atm for c in cols], dev)
        za = lambda: torch.zeros(n * VK_MAX_ATOMS, dtype=F64, device=dev)
        self.atom_ini, self.atom_loss, self.atom_loss_prev = za(), za(), za()
        self.dt = torch.tensor([float(c.dt) for c in cols], dtype=F64, device=dev)
        zc = lambda: torch.zeros(n, dtype=F64, device=dev)
        self.t, self.delta, self.rel = zc(), zc(), zc()
        zi = lambda: torch.zeros(n, dtype=torch.int32, device=dev)
        (self.accept, self.retry, self.nega, self.loss, self.dcount, self.status,
         self.attempts, self.photo_due, self.photo_cad) = (zi() for _ in range(9))
        zero_bot = any(bool(c.atm.use_botflux) for c in cols)
        self.params = cfg.c_params(self.na, zero_bot)
        _lib.check(_lib.load().vk_batch_set_latency_mode(b.handle, int(bool(latency_mode))),
                   "vk_batch_set_latency_mode")
        self.ws = b.workspace()
        self.photo = None
        if photo is not None or any(getattr(c, "photo", None) is not None for c in cols):
            from .photo import PhotoBatch
            self.photo = PhotoBatch(self, cols)
        self.refresh = None
        self._dz_map = None
        if refresh is not None and cfg.update_frq > 0:
            from .atm_refresh import RefreshSet
            self.refresh = RefreshSet(b, refresh)
            if self.photo is not None:
                # the photolysis layer thicknesses follow the refreshed
                # geometry (the reference's photo branch reads the carried
                # dz, outer_loop.py:285, 841)
                st = self.photo.setup
                idx_l = np.concatenate([st.off_l_np[c] + np.arange(int(b.nz[c]))
                                        for c in range(n)])
                idx_g = np.concatenate([self.refresh.off_g_h[c] + 3 * int(b.nz[c]) +
                                        np.arange(int(b.nz[c])) for c in range(n)])
                self._dz_map = (torch.as_tensor(idx_l, device=dev),
                                torch.as_tensor(idx_g, device=dev))
                self.refresh.geo[self._dz_map[1]] = st.dz[self._dz_map[0]]
        self.pin = self.off_pin = None
        if constraints is not None:
            if len(constraints) != b.ncols:
                raise ValueError(f"need one constraints entry (or None) per column ({b.ncols})")
            recs = [(cc or ColumnConstraints()).record(int(b.ni[i]))
                    for i, cc in enumerate(constraints)]
            offs = np.concatenate([[0], np.cumsum([r.size for r in recs])[:-1]]).astype(np.int64)
            self._pin_host = np.concatenate(recs)
            self.pin = torch.tensor(self._pin_host, dtype=F64, device=dev)
            self.off_pin = torch.tensor(offs, dtype=torch.int64, device=dev)
        self.ring_y = self.ring_t = self.conv = None
        self._conv_ptrs = (None,) * 7
        if cfg.conv_mode == "longdy":
            self._setup_conv(conv)
        self.state = SsState(*[t.data_ptr() for t in (
            self.y, self.y_prev, self.ymix, self.sol, self.delta_arr, self.compo, self.atom_ini,
            self.atom_loss, self.atom_loss_prev, self.dt, self.t, self.delta, self.rel,
            self.accept, self.retry, self.nega, self.loss, self.dcount, self.status,
            self.attempts, self.photo_due)], None if self.photo is None
            else self.photo.state.aflux_change.data_ptr(),
            None if self.refresh is None else self.refresh.due.data_ptr(),
            None if self.pin is None else self.pin.data_ptr(),
            None if self.off_pin is None else self.off_pin.data_ptr(), *self._conv_ptrs,
            self.photo_cad.data_ptr())
        self.graph = None
        self._init()

    def _setup_conv(self, conv):
        b, dev, cfg = self.batch, self.dev, self.cfg
        if conv is None or len(conv) != b.ncols:
            raise ValueError(f"conv_mode 'longdy' needs one ColumnConv per column ({b.ncols})")
        if cfg.conv_step < 2:
            raise ValueError("conv_step must be at least 2")
        hp, ign = [], []
        for c, cc in enumerate(conv):
            h = np.asarray(cc.Hp, np.float64).reshape(-1)
            if h.shape[0] != int(b.nz[c]):
                raise ValueError(f"column {c}: Hp has {h.shape[0]} entries, expected {b.nz[c]}")
            m = np.zeros(int(b.ni[c]))
            idx = np.asarray(list(cc.ignore), np.int64)
            if idx.size and (idx.min() < 0 or idx.max() >= b.ni[c]):
                raise ValueError(f"column {c}: conver_ignore index out of range")
            m[idx] = 1.0
            hp.append(h)
            ign.append(m)
        ex = lambda a: np.concatenate([[0], np.cumsum([x.size for x in a])[:-1]]).astype(np.int64)
        self._hp_host = hp
        self.conv_ign = torch.tensor(np.concatenate(ign), dtype=F64, device=dev)
        self.off_ign = torch.tensor(ex(ign), dtype=torch.int64, device=dev)
        if self.refresh is not None:
            # Hp lives in the refresh geometry (mu, g, Hp, dz, zco per column)
            self.hp = self.refresh.geo
            off = self.refresh.off_g_h + 2 * b.nz.astype(np.int64)
            self._hp_idx = torch.as_tensor(np.concatenate(
                [off[c] + np.arange(int(b.nz[c])) for c in range(b.ncols)]), device=dev)
            self.hp[self._hp_idx] = torch.as_tensor(np.concatenate(hp), device=dev)
        else:
            self.hp = torch.tensor(np.concatenate(hp), dtype=F64, device=dev)
            off = ex(hp)
            self._hp_idx = None
        self.off_hp = torch.tensor(off, dtype=torch.int64, device=dev)
        n = int(cfg.conv_step)
        self.ring_y = torch.zeros(n * b.state_len, dtype=F64, device=dev)
        self.ring_t = torch.zeros(b.ncols * n, dtype=F64, device=dev)
        self.conv = torch.zeros(b.ncols * 4, dtype=F64, device=dev)
        self._conv_ptrs = tuple(t.data_ptr() for t in (
            self.ring_y, self.ring_t, self.conv, self.hp, self.off_hp, self.conv_ign,
            self.off_ign))

    def history(self, c: int):
        """Column c's ring in time order, (y_time (n, nz, ni), t_time (n,)),
        n = min(accept_count, conv_step): the reference's ring unpacking."""
        if self.ring_y is None:
            raise RuntimeError("history() needs conv_mode 'longdy'")
        b, n = self.batch, int(self.cfg.conv_step)
        acc = int(self.accept[c].item())
        nz, ni = int(b.nz[c]), int(b.ni[c])
        ry = self.ring_y[n * int(b.off_s[c]):n * int(b.off_s[c]) + n * nz * ni].view(n, nz, ni)
        rt = self.ring_t[c * n:(c + 1) * n]
        start = acc % n if acc >= n else 0
        order = torch.as_tensor([(start + i) % n for i in range(min(acc, n))], device=self.dev)
        return ry[order].cpu().numpy(), rt[order].cpu().numpy()

    def conv_state(self):
        """(ncols, 4): longdy, longdydt, smallest longdy seen, accepted
        steps since it last dropped by 5 %."""
        if self.conv is None:
            raise RuntimeError("conv_state() needs conv_mode 'longdy'")
        return self.conv.view(-1, 4).cpu().numpy()

    # ------------------------------------------------------------------
    def _init(self):
        lib = _lib.load()
        _lib.check(lib.vk_ss_init_f64(self.batch.handle, self.atm.data_ptr(), C.byref(self.params),
                                      C.byref(self.state), _lib.stream_ptr()), "vk_ss_init_f64")

    def attempt(self):
        """One attempt (photo refresh if due, Ros2 step, controller) on the current stream."""
        lib = _lib.load()
        if self.photo is not None:
            self.photo.update()
        _lib.check(lib.vk_ss_attempt_f64(self.batch.handle, self.k.data_ptr(), self.atm.data_ptr(),
                                         C.byref(self.params), C.byref(self.state),
                                         self.ws.data_ptr(), self.ws.numel(), _lib.stream_ptr()),
                   "vk_ss_attempt_f64")
        if self.refresh is not None:
            # the columns the controller flagged: geometry and escape from the
            # clipped step output, read by the next attempt
            self.refresh.run(self.atm, mode=3, ymix=self.ymix, sol=self.sol, params=self.params,
                             use_due=True)
            if self._dz_map is not None:
                st = self.photo.setup
                st.dz[self._dz_map[0]] = self.refresh.geo[self._dz_map[1]]

    def capture(self, n_attempts: int):
        """Capture n attempts in one HIP graph (replayed by run())."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.attempt()          # warm the code paths outside the capture
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(n_attempts):
                self.attempt()
        self.graph = (g, n_attempts)
        self.reset_to_initial()
        return g

    def reset_to_initial(self):
        b = self.batch
        self.y.copy_(b.pack_state([c.y for c in self.columns], self.dev))
        self.dt.copy_(torch.tensor([float(c.dt) for c in self.columns], dtype=F64))
        self.k.copy_(b.pack_k([c.k for c in self.columns], self.dev))
        if self.refresh is not None:
            self.atm.copy_(b.pack_atm([c.atm for c in self.columns], self.dev))
            self.refresh.due.zero_()
            if self._dz_map is not None:
                st = self.photo.setup
                st.dz.copy_(torch.as_tensor(np.concatenate(
                    [np.asarray(c.dz, np.float64) for c in self.columns])))
                self.refresh.geo[self._dz_map[1]] = st.dz[self._dz_map[0]]
        if self.pin is not None:     # the H2/He pin state starts unpinned
            self.pin.copy_(torch.as_tensor(self._pin_host))
        if self.ring_y is not None and self._hp_idx is not None:
            self.hp[self._hp_idx] = torch.as_tensor(np.concatenate(self._hp_host), device=self.dev)
        if self.photo is not None:
            self.photo.reset()
        self._init()

    def done(self) -> bool:
        return bool(((self.status & ST_DONE) != 0).all().item())

    def run(self, max_attempts: int = 100000, check_every: int = 50, use_graph: bool = False,
            report_every: float = 0.0, metrics=None):
        """Attempt until every column is done or max_attempts; returns attempts made.
        report_every > 0 prints progress after the check that crosses each interval (s).
        ``metrics`` (a path, stream or metrics.JsonlLog) receives a
        ``steady_check`` record at every check."""
        import time
        from .metrics import open_log
        log = open_log(metrics)
        t0 = time.perf_counter()
        next_report = report_every
        done_attempts = 0
        if use_graph and self.graph is None:
            self.capture(check_every)
        while done_attempts < max_attempts:
            if use_graph:
                self.graph[0].replay()
                done_attempts += self.graph[1]
            else:
                for _ in range(check_every):
                    self.attempt()
                done_attempts += check_every
            finished = self.done()
            if log is not None:
                self._log_check(log, done_attempts, time.perf_counter() - t0)
            if finished:
                break
            if report_every > 0 and time.perf_counter() - t0 >= next_report:
                next_report += report_every
                n_done = int(((self.status & ST_DONE) != 0).sum().item())
                print(f"[steady] t={time.perf_counter() - t0:.0f}s attempts={done_attempts} "
                      f"done={n_done}/{self.batch.ncols}", flush=True)
        torch.cuda.synchronize()
        return done_attempts

    def _log_check(self, log, attempts_done: int, wall: float):
        st = self.status.cpu().numpy()
        acc = self.accept.cpu().numpy()
        att = self.attempts.cpu().numpy()
        log.write("steady_check", columns=int(st.size), attempts=int(attempts_done),
                  wall_s=round(wall, 6), done=int(((st & ST_DONE) != 0).sum()),
                  converged=int(((st & ST_CONVERGED) != 0).sum()),
                  capped=int(((st & ST_MAXSTEPS) != 0).sum()),
                  accepted_total=int(acc.sum()), attempts_total=int(att.sum()),
                  column_attempts_per_s=round(float(att.sum()) / wall, 3) if wall > 0 else None,
                  rejects={"delta": int(self.dcount.sum().item()),
                           "nega": int(self.nega.sum().item()),
                           "loss": int(self.loss.sum().item())})

    # ------------------------------------------------------------------
    # checkpoint / resume (the run state the reference packs and unpacks,
    # outer_loop.py:1795-2013)
    _STATE = ("y", "y_prev", "ymix", "sol", "delta_arr", "atom_ini", "atom_loss",
              "atom_loss_prev", "dt", "t", "delta", "rel", "accept", "retry", "nega", "loss",
              "dcount", "status", "attempts", "photo_due", "photo_cad", "k", "atm")

    def snapshot(self) -> dict:
        """Every device array the next attempt reads, copied to the host
        (NumPy), after the stream's pending work."""
        torch.cuda.current_stream().synchronize()
        snap = {f: getattr(self, f).detach().cpu().numpy().copy() for f in self._STATE}
        if self.refresh is not None:
            snap["refresh_due"] = self.refresh.due.cpu().numpy().copy()
            snap["refresh_geo"] = self.refresh.geo.cpu().numpy().copy()
        if self.photo is not None:
            for name, t in self.photo.state_tensors().items():
                snap["photo_" + name] = t.detach().cpu().numpy().copy()
        if self.pin is not None:
            snap["pin"] = self.pin.cpu().numpy().copy()
        if self._dz_map is not None:
            snap["photo_dz"] = self.photo.setup.dz.cpu().numpy().copy()
        if self.ring_y is not None:
            for f in ("ring_y", "ring_t", "conv"):
                snap[f] = getattr(self, f).cpu().numpy().copy()
        return snap

    def restore(self, snap: dict):
        """Load a snapshot of a runner built on the same columns; the next
        attempt continues exactly where the snapshot was taken."""
        for f in self._STATE:
            t = getattr(self, f)
            a = np.asarray(snap[f])
            if a.shape != tuple(t.shape):
                raise ValueError(f"snapshot field {f} has shape {a.shape}, expected "
                                 f"{tuple(t.shape)}")
            t.copy_(torch.as_tensor(a))
        if self.refresh is not None:
            self.refresh.due.copy_(torch.as_tensor(snap["refresh_due"]))
            self.refresh.geo.copy_(torch.as_tensor(snap["refresh_geo"]))
        if self.photo is not None:
            for name, t in self.photo.state_tensors().items():
                t.copy_(torch.as_tensor(snap["photo_" + name]))
        if self.pin is not None:
            self.pin.copy_(torch.as_tensor(snap["pin"]))
        if self._dz_map is not None:
            self.photo.setup.dz.copy_(torch.as_tensor(snap["photo_dz"]))
        if self.ring_y is not None:
            for f in ("ring_y", "ring_t", "conv"):
                getattr(self, f).copy_(torch.as_tensor(snap[f]))

    def save_checkpoint(self, path):
        """snapshot() into an .npz file (no pickled objects)."""
        np.savez(path, **self.snapshot())

    def load_checkpoint(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.restore({k: z[k] for k in z.files})

    def results(self):
        b = self.batch
        ys = [t.detach().cpu().numpy() for t in b.unpack_state(self.y)]
        mixes = [t.detach().cpu().numpy() for t in b.unpack_state(self.ymix)]
        return dict(
            y=ys, ymix=mixes, status=self.status.cpu().numpy(), accept=self.accept.cpu().numpy(),
            attempts=self.attempts.cpu().numpy(), t=self.t.cpu().numpy(), dt=self.dt.cpu().numpy(),
            rel=self.rel.cpu().numpy(), delta=self.delta.cpu().numpy(),
            counts=dict(nega=self.nega.cpu().numpy(), loss=self.loss.cpu().numpy(),
                        delta=self.dcount.cpu().numpy()))
