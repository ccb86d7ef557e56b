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
