"""GPU: the allocation-free Ros2 step (vk_ros2_step_noalloc_f64).

It runs the kernels of vk_ros2_step_f64 on a file-scope device scratch
array instead of a caller workspace.  Two calls in a row, the second with a
smaller batch (nz = 150 then nz = 40, ni = 93), each equal bit for bit to
vk_ros2_step_f64 on the same inputs; then the same for one column in
latency mode (cyclic reduction, the largest workspace per column); a batch
whose workspace exceeds the array is refused with a clear error and
launches nothing.  Every call goes through the C ABI.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

NI, NREV, NPHOTO = 93, 1200, 60
SCRATCH = 160 << 20          # VK_NOALLOC_SCRATCH_BYTES
VK_ERR_UNSUPPORTED = 4


def _inputs(seed, B, nz):
    import torch
    from vulcan_amd import core, synth
    cols = synth.columns_same_network(seed, B, nz, NI, NREV, NPHOTO, dt=1e-10, name="na")
    dev = torch.device("cuda", 0)
    y = torch.tensor(np.stack([c.y for c in cols]), dtype=torch.float64, device=dev)
    k = torch.tensor(np.stack([c.k for c in cols]), dtype=torch.float64, device=dev)
    atm = torch.cat([core.pack_atm_record(c.atm) for c in cols]).to(dev)
    dt = torch.full((B,), 1e-10, dtype=torch.float64, device=dev)
    return cols[0].net, y, k, atm, dt


def _both(b, y, k, atm, dt):
    """(sol, delta, status) of the workspace step and of the allocation-free one."""
    import torch
    from vulcan_amd import _lib
    lib = _lib.load()
    out = []
    for noalloc in (False, True):
        sol = torch.full_like(y, float("nan"))
        delta = torch.full_like(y, float("nan"))
        status = torch.zeros(y.shape[0], dtype=torch.int32, device=y.device)
        if noalloc:
            rc = lib.vk_ros2_step_noalloc_f64(b.handle, y.data_ptr(), k.data_ptr(),
                                              atm.data_ptr(), dt.data_ptr(), None,
                                              sol.data_ptr(), delta.data_ptr(),
                                              status.data_ptr(), _lib.stream_ptr())
        else:
            ws = b.workspace()
            rc = lib.vk_ros2_step_f64(b.handle, y.data_ptr(), k.data_ptr(), atm.data_ptr(),
                                      dt.data_ptr(), None, sol.data_ptr(), delta.data_ptr(),
                                      ws.data_ptr(), ws.numel(), status.data_ptr(),
                                      _lib.stream_ptr())
        _lib.check(rc, "noalloc step" if noalloc else "workspace step")
        torch.cuda.synchronize()
        out.append((sol, delta, status))
    return out


def test_noalloc_step_equals_workspace_step():
    import torch
    from vulcan_amd import _lib, core
    lib = _lib.load()
    for seed, B, nz in ((1201, 3, 150), (1202, 2, 40)):     # the second call is smaller
        net, y, k, atm, dt = _inputs(seed, B, nz)
        b = core.uniform_batch(B, nz, net)
        assert lib.vk_ros2_workspace_bytes(b.handle) <= SCRATCH
        (s0, d0, st0), (s1, d1, st1) = _both(b, y, k, atm, dt)
        print(f"B={B} nz={nz}: workspace {lib.vk_ros2_workspace_bytes(b.handle) / 2**20:.1f} MiB, "
              f"max |sol diff| {float((s0 - s1).abs().max()):.3e}")
        assert torch.isfinite(s0).all()
        assert torch.equal(s0, s1) and torch.equal(d0, d1) and torch.equal(st0, st1)


def test_noalloc_step_latency_mode_and_size_limit():
    import torch
    from vulcan_amd import _lib, core
    lib = _lib.load()
    net, y, k, atm, dt = _inputs(1203, 1, 150)
    b = core.Batch([150], [core.device_network(net)], [0])
    _lib.check(lib.vk_batch_set_latency_mode(b.handle, 1), "vk_batch_set_latency_mode")
    need = lib.vk_ros2_workspace_bytes(b.handle)
    assert need <= SCRATCH
    (s0, d0, st0), (s1, d1, st1) = _both(b, y, k, atm, dt)
    print(f"latency mode, one column nz=150: workspace {need / 2**20:.1f} MiB, equal "
          f"{torch.equal(s0, s1)}")
    assert torch.isfinite(s0).all() and torch.equal(s0, s1) and torch.equal(d0, d1)
    # eight columns of nz=150 need more than the array: refused before any launch
    net8, y8, k8, atm8, dt8 = _inputs(1204, 8, 150)
    b8 = core.uniform_batch(8, 150, net8)
    assert lib.vk_ros2_workspace_bytes(b8.handle) > SCRATCH
    sol = torch.full_like(y8, 7.0)
    delta = torch.full_like(y8, 7.0)
    status = torch.zeros(8, dtype=torch.int32, device=y8.device)
    rc = lib.vk_ros2_step_noalloc_f64(b8.handle, y8.data_ptr(), k8.data_ptr(), atm8.data_ptr(),
                                      dt8.data_ptr(), None, sol.data_ptr(), delta.data_ptr(),
                                      status.data_ptr(), _lib.stream_ptr())
    torch.cuda.synchronize()
    msg = lib.vk_last_error().decode()
    print(f"8 columns: rc {rc}: {msg}")
    assert rc == VK_ERR_UNSUPPORTED and "scratch" in msg
    assert bool((sol == 7.0).all()) and bool((delta == 7.0).all())
