"""The C-ABI library loads without a GPU and exports every entry point include/woo_amd.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "woo_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\*\s]+?\b(wdg_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "wdg_collide" in syms and "wdg_contact_loop" in syms and "wdg_leapfrog" in syms
    from woo_amd._native import EXPORTS

    assert sorted(EXPORTS) == syms


def test_library_exports_all_declared_symbols():
    from woo_amd._native import LIB_PATH

    if not os.path.exists(LIB_PATH):
        pytest.skip("libwoo_amd.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import woo_amd._native as nat

    monkeypatch.setattr(nat, "_lib", None)
    monkeypatch.setattr(nat, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(ImportError):
        nat.lib()


def test_error_path_without_device():
    """wdg_ctx_create on a machine without a HIP device returns an error code and a message."""
    from woo_amd._native import LIB_PATH

    if not os.path.exists(LIB_PATH):
        pytest.skip("libwoo_amd.so not built")
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from woo_amd._native import lib

    L = lib()
    h = ctypes.c_void_p()
    rc = L.wdg_ctx_create(0, ctypes.byref(h))
    assert rc < 0
    assert len(L.wdg_last_error()) > 0
