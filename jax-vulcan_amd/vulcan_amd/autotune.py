"""Launch-shape autotuner for batches of uneven columns.

For every 16-species bucket (SL = ceil(ni/16)) present in a batch, time the
block-factor kernels the library offers for that bucket (``VK_FACTOR_TILE``:
one 256-thread workgroup per column with 16x16 register tiles;
``VK_FACTOR_ROWS``: half a row per lane, 64-192 lanes per column;
``VK_FACTOR_BLOCKED``: 16-column panels with FP64 MFMA trailing updates;
``VK_FACTOR_TILE2`` / ``VK_FACTOR_WAVE`` / ``VK_FACTOR_TILE8``: the tile
kernel on two / one / eight waves per column; ``VK_FACTOR_CR``: block cyclic
reduction, log2(nz) levels of one workgroup per block row, the latency mode
for a few columns) on
the batch's own columns of that bucket, and keep the fastest
(``vk_tune_set``).  All take the same pivots; TILE, ROWS, TILE2, WAVE and
TILE8 give bitwise-equal inverses, BLOCKED differs in the rounding of its rank-16
products, CR in the order its rows are combined.  Decisions are cached per
(SL, nz rounded up to 32, column count rounded up to a power of two).

The timed call is one plain Ros2 step of the bucket's columns (y -> sol), so
a runner's state is not touched.  With several buckets, whose kernels share
the GPU inside one step (forked streams), the two fastest variants of each
bucket alone are then combined and the whole batch's step is timed for
every combination: a bucket's best alone is not always its best beside the
others (the cyclic-reduction factor does twice the tile factor's work, which
pays when a bucket is alone on the chip).  Tune before building runners or serving
pools: the cyclic-reduction factor needs a larger step workspace
(vk_ros2_workspace_bytes), sized when they are built.
"""
from __future__ import annotations

import ctypes as C
import itertools
import math
import time
from typing import Dict, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .core import F64, Batch, device, device_network

VK_TUNE_FACTOR = 0
VARIANTS = {0: "tile", 1: "rows", 2: "blocked", 3: "tile2", 4: "wave", 5: "tile8", 6: "cr"}
_CACHE: Dict[Tuple[int, int, int], int] = {}
_TIMES: Dict[Tuple[int, int, int], Dict[int, float]] = {}
_JOINT: Dict[tuple, tuple] = {}


def _key(sl: int, nz_max: int, ncols: int) -> Tuple[int, int, int]:
    return (sl, 32 * math.ceil(nz_max / 32), 1 << max(0, math.ceil(math.log2(max(ncols, 1)))))


def _time_step(batch: Batch, y, k, atm, dt, reps: int) -> float:
    lib = _lib.load()
    ws = batch.workspace()
    sol = torch.empty_like(y)
    delta = torch.empty_like(y)
    status = torch.zeros(batch.ncols, dtype=torch.int32, device=y.device)
    st = torch.cuda.current_stream().cuda_stream

    def run():
        _lib.check(lib.vk_ros2_step_f64(batch.handle, y.data_ptr(), k.data_ptr(), atm.data_ptr(),
                                        dt.data_ptr(), None, sol.data_ptr(), delta.data_ptr(),
                                        ws.data_ptr(), ws.numel(), status.data_ptr(), st),
                   "vk_ros2_step_f64")

    run()  # first launch: module load, LDS attributes
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def _step_inputs(cols, dev):
    nets, ids, seen = [], [], {}
    for c in cols:
        if id(c.net) not in seen:
            seen[id(c.net)] = len(nets)
            nets.append(device_network(c.net))
        ids.append(seen[id(c.net)])
    b = Batch([c.nz for c in cols], nets, ids)
    y = b.pack_state([c.y for c in cols], dev)
    k = b.pack_k([c.k for c in cols], dev)
    atm = b.pack_atm([c.atm for c in cols], dev)
    dt = torch.tensor([float(c.dt) for c in cols], dtype=F64, device=dev)
    return b, y, k, atm, dt


def tune_columns(columns: Sequence, reps: int = 2, verbose: bool = False,
                 joint: bool = True) -> Dict[int, str]:
    """Pick the block-factor kernel for each bucket of ``columns``; returns
    {SL: variant name}.  Cached decisions are applied without timing.
    ``joint`` (default) re-times the two fastest variants of each bucket
    together when there are several buckets."""
    lib = _lib.load()
    dev = device()
    by_sl: Dict[int, list] = {}
    for c in columns:
        by_sl.setdefault((int(c.ni) + 15) // 16, []).append(c)
    chosen = {}
    keys = {}
    for sl, cols in sorted(by_sl.items()):
        key = _key(sl, max(int(c.nz) for c in cols), len(cols))
        keys[sl] = key
        if key in _CACHE:
            _lib.check(lib.vk_tune_set(VK_TUNE_FACTOR, sl, _CACHE[key]), "vk_tune_set")
            chosen[sl] = VARIANTS[_CACHE[key]]
            continue
        b, y, k, atm, dt = _step_inputs(cols, dev)
        times = {}
        for v in VARIANTS:
            if lib.vk_tune_set(VK_TUNE_FACTOR, sl, v) != 0:
                continue  # variant not offered for this bucket
            times[v] = _time_step(b, y, k, atm, dt, reps)
        best = min(times, key=times.get)
        _lib.check(lib.vk_tune_set(VK_TUNE_FACTOR, sl, best), "vk_tune_set")
        _CACHE[key] = best
        _TIMES[key] = times
        chosen[sl] = VARIANTS[best]
        if verbose:
            print(f"[autotune] SL={sl} ({len(cols)} columns, nz<={key[1]}): " +
                  ", ".join(f"{VARIANTS[v]} {t * 1e3:.2f} ms" for v, t in times.items()) +
                  f" -> {VARIANTS[best]}", flush=True)
    if joint and len(by_sl) > 1 and all(keys[sl] in _TIMES for sl in by_sl):
        sls = sorted(by_sl)
        jkey = tuple(keys[sl] for sl in sls)
        if jkey not in _JOINT:
            cands = [sorted(_TIMES[keys[sl]], key=_TIMES[keys[sl]].get)[:2] for sl in sls]
            b, y, k, atm, dt = _step_inputs(list(columns), dev)
            jt = {}
            for combo in itertools.product(*cands):
                for sl, v in zip(sls, combo):
                    _lib.check(lib.vk_tune_set(VK_TUNE_FACTOR, sl, v), "vk_tune_set")
                jt[combo] = _time_step(b, y, k, atm, dt, reps)
            _JOINT[jkey] = min(jt, key=jt.get)
            if verbose:
                print("[autotune] buckets together: " + ", ".join(
                    "/".join(f"{sl}:{VARIANTS[v]}" for sl, v in zip(sls, cb)) +
                    f" {t * 1e3:.2f} ms" for cb, t in jt.items()) + " -> " +
                    "/".join(f"{sl}:{VARIANTS[v]}" for sl, v in zip(sls, _JOINT[jkey])),
                    flush=True)
        for sl, v in zip(sls, _JOINT[jkey]):
            _lib.check(lib.vk_tune_set(VK_TUNE_FACTOR, sl, v), "vk_tune_set")
            chosen[sl] = VARIANTS[v]
    return chosen


def clear_cache():
    _CACHE.clear()
    _TIMES.clear()
    _JOINT.clear()
