"""The C-ABI library loads and exports every symbol include/*.h declares."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(vk_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_all_declared_symbols():
    import ctypes
    from vulcan_amd import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_symbols()) >= 20


def test_ctypes_signatures_cover_header():
    from vulcan_amd import _lib
    assert declared_symbols() <= set(_lib.exported_symbols())


def test_host_calls_without_gpu():
    from vulcan_amd import _lib
    lib = _lib.load()
    assert lib.vk_version() >= 100
    assert lib.vk_max_species() == 112
    assert lib.vk_block_doubles(93) == 6 * 6 * 256
    assert lib.vk_block_doubles(50) == 4 * 4 * 256


def test_atm_record_layout_matches_python():
    from vulcan_amd import core
    total, off = core.atm_record_layout(10, 7)
    nz, ni = 10, 7
    assert off[0] == 0 and off[5] == 5 * (nz - 1)
    assert total == 5 * (nz - 1) + 3 * nz + 6 * ni + 2 * (nz - 1) * ni + nz * ni + 4


def test_rates_create_rejects_bad_tables_without_gpu():
    """vk_rates_create validates its tables on the host before any device
    call: an unknown form code or an out-of-range species index is a bad
    argument."""
    import ctypes as C
    import numpy as np
    from vulcan_amd import _lib
    lib = _lib.load()
    nr, ni, T = 2, 3, 2
    form = np.array([0, 1, 0], np.int32)
    z = np.zeros(nr + 1, np.int32)
    par = np.zeros((nr + 1, 6))
    idx = np.full((nr + 1, T), ni, np.int64)
    st = np.zeros((nr + 1, T))
    h = C.c_void_p()

    def create(form, idx):
        return lib.vk_rates_create(nr, ni, T, nr + 1, form.ctypes.data, z.ctypes.data,
                                   z.ctypes.data, par.ctypes.data, idx.ctypes.data,
                                   st.ctypes.data, idx.ctypes.data, st.ctypes.data, None, None,
                                   C.byref(h))

    bad_form = form.copy()
    bad_form[1] = 7
    assert create(bad_form, idx) == 1
    assert b"form" in lib.vk_last_error()
    bad_idx = idx.copy()
    bad_idx[1, 0] = ni + 1
    assert create(form, bad_idx) == 1
    assert lib.vk_rates_create(nr, 0, T, 1, None, None, None, None, None, None, None, None,
                               None, None, C.byref(h)) == 1


def test_ctypes_structs_match_the_header(tmp_path):
    """Every structure the Python side mirrors has the header's size and
    field offsets (compiled with the host C compiler)."""
    import shutil
    import subprocess
    from vulcan_amd import atm_refresh, outer_loop, photo
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    mirrors = {"vk_ss_params": outer_loop.SsParams, "vk_ss_state": outer_loop.SsState,
               "vk_refresh_state": atm_refresh.RefreshState,
               "vk_photo_setup": photo.PhotoSetupC, "vk_photo_state": photo.PhotoStateC}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "vulcan_ros2.h"',
             'int main(void) {']
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True,
                               text=True).stdout.splitlines():
        cname, f, v = line.split()
        got[(cname, f)] = int(v)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_network_create_rejects_bad_tables_without_gpu():
    """vk_network_create validates sizes, indices and stoichiometry before it
    touches the device: each bad table is refused with its own message."""
    import ctypes as C
    import numpy as np
    from vulcan_amd import _lib
    lib = _lib.load()

    def create(ni, r_idx, p_idx, r_st, p_st, T=2):
        nr = r_idx.shape[0] - 1
        out = C.c_void_p()
        args = [np.ascontiguousarray(a) for a in (r_idx, p_idx, r_st, p_st)]
        rc = lib.vk_network_create(ni, nr, T, *(a.ctypes.data for a in args), None, C.byref(out))
        return rc, lib.vk_last_error().decode()

    ni = 4
    ok_r = np.array([[ni, ni], [0, 1], [2, 3]], np.int64)
    ok_p = np.array([[ni, ni], [2, 3], [0, 1]], np.int64)
    ok_s = np.array([[0.0, 0.0], [1.0, 1.0], [1.0, 1.0]])
    bad_idx = ok_r.copy()
    bad_idx[1, 0] = ni + 1
    rc, msg = create(ni, bad_idx, ok_p, ok_s, ok_s)
    assert rc != 0 and "out of range" in msg
    frac = ok_s.copy()
    frac[2, 1] = 0.5
    rc, msg = create(ni, ok_r, ok_p, frac, ok_s)
    assert rc != 0 and "stoichiometry" in msg
    rc, msg = create(lib.vk_max_species() + 1, ok_r, ok_p, ok_s, ok_s)
    assert rc != 0 and "unsupported sizes" in msg
    rc, msg = create(ni, ok_r, ok_p, ok_s, ok_s, T=5)
    assert rc != 0 and "unsupported sizes" in msg


def test_tuning_knobs_without_gpu():
    """vk_tune_set / vk_tune_get are host-only: defaults, round trips and
    rejected values for the sweep, refinement and layout knobs."""
    from vulcan_amd import _lib
    lib = _lib.load()
    VK_TUNE_REFINE, VK_TUNE_SPLIT, VK_TUNE_KDIRECT, VK_TUNE_HOLDX = 2, 3, 4, 5
    assert lib.vk_tune_get(VK_TUNE_HOLDX, 0) == 1
    assert lib.vk_tune_get(VK_TUNE_REFINE, 0) == 2
    assert lib.vk_tune_get(VK_TUNE_KDIRECT, 0) == 0
    assert lib.vk_tune_get(VK_TUNE_SPLIT, 0) == 1
    try:
        assert lib.vk_tune_set(VK_TUNE_HOLDX, 0, 0) == 0
        assert lib.vk_tune_get(VK_TUNE_HOLDX, 0) == 0
        assert lib.vk_tune_set(VK_TUNE_HOLDX, 0, 2) != 0
        assert b"VK_TUNE_HOLDX" in lib.vk_last_error()
        assert lib.vk_tune_get(VK_TUNE_HOLDX, 0) == 0
    finally:
        lib.vk_tune_set(VK_TUNE_HOLDX, 0, 1)
    assert lib.vk_tune_get(VK_TUNE_HOLDX, 0) == 1
    assert lib.vk_tune_get(99, 0) == -1
