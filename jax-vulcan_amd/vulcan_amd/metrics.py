"""JSON Lines metrics for runs, serving and benchmarks.

One JSON object per line, each with ``event`` and ``t_s`` (seconds since the
log was opened).  The steady-state runner writes a ``steady_check`` record
at each host check, the serving engine a ``request_done`` record per
finished request and a ``serving_summary`` at the end, and ``bench.py
--metrics PATH`` its legs and its result line.  The reference prints
progress only (legacy_io.py:623-673); these records carry the same counters
(accepted steps, the delta / nega / loss rejects, outer_loop.py:58-66) in a
form a collector can read.
"""
from __future__ import annotations

import io
import json
import time
from pathlib import Path
from typing import Any, Optional, Union

import numpy as np

__all__ = ["JsonlLog", "open_log"]


def _plain(v: Any):
    """NumPy scalars and arrays as JSON values."""
    if isinstance(v, np.generic):
        return v.item()
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, dict):
        return {str(k): _plain(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_plain(x) for x in v]
    return v


class JsonlLog:
    """Append JSON records, one per line, to a path or an open text stream."""

    def __init__(self, target: Union[str, Path, io.TextIOBase]):
        self._own = not hasattr(target, "write")
        self._f = open(target, "a", encoding="utf-8") if self._own else target
        self._t0 = time.perf_counter()

    def write(self, event: str, **fields) -> dict:
        rec = {"event": event, "t_s": round(time.perf_counter() - self._t0, 6)}
        rec.update({k: _plain(v) for k, v in fields.items()})
        self._f.write(json.dumps(rec, allow_nan=False, default=str) + "\n")
        self._f.flush()
        return rec

    def close(self):
        if self._own:
            self._f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def open_log(target: Optional[Union[str, Path, io.TextIOBase, JsonlLog]]) -> Optional[JsonlLog]:
    """A JsonlLog for a path or stream; an existing log is returned as is."""
    if target is None or isinstance(target, JsonlLog):
        return target
    return JsonlLog(target)
