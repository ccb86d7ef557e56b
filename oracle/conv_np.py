"""Oracle: the reference's steady-state stopping rule, restated in NumPy.

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module; the
tests, the smoke check and the CPU baseline of the benchmark use it as the
checker of the device controller's conv_mode 1.

Follows outer_loop.py of the reference:

* ``lookback_index``  - the ring entry nearest t * st_factor with the slot
  just written excluded (_conv_jax, outer_loop.py:645-654);
* ``longdy``          - max over layers and species of
  |y - y_old| / n0 / ymix with the mtol_conv, atol and conver_ignore masks
  (outer_loop.py:656-670); condensation masks are not modelled;
* ``slope_min``       - min(Kzz / (0.1 Hp[:-1])^2) clamped to [1e-10, 1e-8]
  (outer_loop.py:676-683);
* ``ConvTracker``     - the ring append on every attempt (slot
  (accept_count_next - 1) % conv_step, outer_loop.py:918-927), longdy and
  longdydt refreshed on accepted steps (:967-979), the stall bookkeeping
  with a 5 % drop as a new minimum (:985-1003) and cond_fn's decision
  (:672-713).

Pins (tests/test_oracle_conv.py): closed forms (a state twice the ring's
gives longdy 0.5 exactly; equal states give 0), a hand-built ring for the
lookback, the clamp's three regimes, a scripted longdy sequence for the
stall counter, and the reference's own count_max run property (the run
stops at accept_count == count_max + 1 with the ring in time order,
tests/test_outer_loop_conv.py:1-25).
"""
from __future__ import annotations

import numpy as np


def lookback_index(t_ring, t, st_factor, accept_count_after):
    """Index of the ring entry nearest t * st_factor, never the slot of the
    step just accepted; ties go to the lowest index (argmin)."""
    t_ring = np.asarray(t_ring, dtype=np.float64)
    n = t_ring.shape[0]
    diffs = np.abs(t_ring - t * st_factor)
    diffs[max(int(accept_count_after) - 1, 0) % n] = np.inf
    return int(np.argmin(diffs))


def longdy(y, y_old, ymix, n0, atol, mtol_conv, ignore=None):
    """max |y - y_old| / n0 / ymix over the unmasked entries."""
    y = np.asarray(y, np.float64)
    d = np.abs((y - y_old) / np.asarray(n0, np.float64)[:, None])
    d = np.where(ymix < mtol_conv, 0.0, d)
    d = np.where(y < atol, 0.0, d)
    if ignore is not None:
        d = np.where(np.asarray(ignore, bool)[None, :], 0.0, d)
    ratio = np.where(ymix > 0, d / np.maximum(ymix, 1e-300), 0.0)
    return float(ratio.max())


def slope_min(Kzz, Hp):
    s = float(np.min(np.asarray(Kzz) / (0.1 * np.asarray(Hp)[:-1]) ** 2))
    return max(min(s, 1e-8), 1e-10)


class ConvTracker:
    """Ring, longdy / longdydt and stall state of one column."""

    def __init__(self, nz, ni, conv_step, st_factor=0.5, yconv_cri=0.01, slope_cri=1e-4,
                 yconv_min=0.1, mtol_conv=1e-20, atol=0.1, conv_stall_window=200,
                 flux_cri=0.1, trun_min=1e2, count_min=120, count_max=5000, runtime=1e22):
        self.conv_step = int(conv_step)
        self.ring_y = np.zeros((conv_step, nz, ni))
        self.ring_t = np.zeros(conv_step)
        self.longdy = 1.0      # state.py:706-707
        self.longdydt = 1.0
        self.seen_min = np.inf
        self.since = 0
        self.p = dict(st_factor=st_factor, yconv_cri=yconv_cri, slope_cri=slope_cri,
                      yconv_min=yconv_min, mtol_conv=mtol_conv, atol=atol,
                      conv_stall_window=conv_stall_window, flux_cri=flux_cri,
                      trun_min=trun_min, count_min=count_min, count_max=count_max,
                      runtime=runtime)

    def update(self, y_next, ymix_next, t_next, accept_count_next, do_accept, n0,
               ignore=None):
        """One attempt of body_fn after the state update."""
        p = self.p
        slot = max(int(accept_count_next) - 1, 0) % self.conv_step
        self.ring_y[slot] = y_next
        self.ring_t[slot] = t_next
        if not do_accept:
            return
        idx = lookback_index(self.ring_t, t_next, p["st_factor"], accept_count_next)
        ld = longdy(y_next, self.ring_y[idx], ymix_next, n0, p["atol"], p["mtol_conv"], ignore)
        self.longdy = ld
        self.longdydt = ld / max(t_next - self.ring_t[idx], 1e-300)
        if ld < self.seen_min * 0.95:
            self.seen_min = ld
            self.since = 0
        else:
            self.since += 1

    def terminate(self, t, accept_count, Kzz, Hp, aflux_change=0.0):
        """cond_fn's decision: (converged_or_stalled_and_ready, too_long_or_too_many)."""
        p = self.p
        sm = slope_min(Kzz, Hp)
        flux_ok = aflux_change < p["flux_cri"]
        conv = (((self.longdy < p["yconv_cri"]) and (self.longdydt < p["slope_cri"])) or
                ((self.longdy < p["yconv_min"]) and (self.longdydt < sm))) and flux_ok
        stalled = (self.since > p["conv_stall_window"] and self.seen_min < p["yconv_min"] and
                   self.longdy < p["yconv_min"] and flux_ok)
        ready = t > p["trun_min"] and accept_count > p["count_min"]
        return (ready and (conv or stalled),
                t > p["runtime"] or accept_count > p["count_max"])

    def history(self, accept_count):
        """(y_time, t_time) in time order: the last min(count, conv_step)
        accepted states (the reference's ring unpacking)."""
        n = min(int(accept_count), self.conv_step)
        start = int(accept_count) % self.conv_step if accept_count >= self.conv_step else 0
        order = [(start + i) % self.conv_step for i in range(n)]
        return self.ring_y[order], self.ring_t[order]
