"""The C ABI library loads without a GPU and exports every symbol the public headers declare."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        pat = r"^\s*(?:const\s+)?(?:int|void|long|double|float|char|unsigned|size_t|tti_\w+)\s*\*?\s*(\w+)\s*\("
        for m in re.finditer(pat, text, flags=re.M):
            names.append(m.group(1))
    return names


def test_headers_declare_entry_points():
    names = declared_symbols()
    assert len(names) >= 12
    assert "launch_fwd_kernel2_drv" in names and "launch_update_rho_fwd_main_loop_2_1" in names


def test_library_exports_all_declared_symbols():
    from miniappsforhpe_amd import _lib

    if not _lib.LIB_PATH.exists():
        pytest.fail("libtti_hip.so is not built (run python -m miniappsforhpe_amd.build)")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"symbols declared in include/*.h but not exported: {missing}"


def test_binding_covers_every_declared_symbol():
    from miniappsforhpe_amd import _lib

    bound = set(_lib._SIGS) | set(_lib._LONG_RET) | set(_lib._OTHER_RET)
    missing = [n for n in declared_symbols() if n not in bound]
    assert not missing, f"no ctypes signature for {missing}"
