"""CPU: host-side validation of the run configuration (no GPU calls)."""
from __future__ import annotations

import pytest


def test_conv_mode_is_validated():
    from vulcan_amd.outer_loop import RunConfig
    p = RunConfig(conv_mode="longdy", conv_step=40, yconv_cri=0.02).c_params(1)
    assert p.conv_mode == 1 and p.conv_step == 40 and p.yconv_cri == 0.02
    assert p.st_factor == 0.5 and p.conv_stall_window == 200 and p.runtime == 1e22
    assert RunConfig().c_params(1).conv_mode == 0
    with pytest.raises(ValueError):
        RunConfig(conv_mode="window").c_params(1)


def test_serving_refuses_the_longdy_rule():
    from vulcan_amd.outer_loop import RunConfig
    from vulcan_amd.serving import ServingEngine
    with pytest.raises(ValueError, match="relmax"):
        ServingEngine(cfg=RunConfig(conv_mode="longdy"))


def test_vul_end_case_codes():
    """The .vul end_case follows the reference (outer_loop.py:1938-1942,
    2326-2339): 3 when the step cap was exceeded, 2 when the runtime was,
    1 otherwise; never the device status word."""
    from vulcan_amd.output import end_case
    from vulcan_amd.outer_loop import RunConfig, ST_CONVERGED, ST_MAXSTEPS
    ld = RunConfig(conv_mode="longdy", count_max=100, runtime=1e5)
    assert end_case(101, 1.0, ST_MAXSTEPS, ld) == 3
    assert end_case(100, 1.0, 0, ld) == 1
    assert end_case(50, 2e5, ST_MAXSTEPS, ld) == 2
    assert end_case(50, 1.0, ST_CONVERGED, ld) == 1
    rm = RunConfig(count_max=100)
    assert end_case(100, 1.0, ST_MAXSTEPS, rm) == 3
    assert end_case(100, 1.0, ST_MAXSTEPS | ST_CONVERGED, rm) == 1
    assert end_case(100, 1.0, ST_CONVERGED | 2, rm) == 1
    assert end_case(40, 1.0, 0, rm) == 1
