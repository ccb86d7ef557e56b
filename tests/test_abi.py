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
