"""The C-ABI library loads and exports every symbol include/*.h declares."""
from __future__ import annotations

import re
from pathlib import Path

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
