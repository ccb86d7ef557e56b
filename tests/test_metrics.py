"""CPU: the JSON Lines metrics log."""
from __future__ import annotations

import io
import json

import numpy as np


def test_jsonl_records_are_one_json_object_per_line(tmp_path):
    from vulcan_amd.metrics import JsonlLog, open_log
    path = tmp_path / "m.jsonl"
    with JsonlLog(path) as log:
        log.write("steady_check", attempts=np.int64(50), rejects={"delta": np.int32(3)},
                  ys=np.array([1.0, 2.0]))
        log.write("serving_summary", p50_s=1.5, throughput_solves_per_s=None)
    lines = path.read_text().splitlines()
    assert len(lines) == 2
    a, b = (json.loads(line) for line in lines)
    assert a["event"] == "steady_check" and a["attempts"] == 50 and a["rejects"] == {"delta": 3}
    assert a["ys"] == [1.0, 2.0] and a["t_s"] >= 0.0
    assert b["event"] == "serving_summary" and b["throughput_solves_per_s"] is None
    # appending to an open stream, and passing a log through
    buf = io.StringIO()
    log = open_log(buf)
    assert open_log(log) is log and open_log(None) is None
    log.write("bench_result", value=1.0)
    assert json.loads(buf.getvalue())["value"] == 1.0
