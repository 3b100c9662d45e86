"""The C-ABI library loads (no GPU needed) and exports every entry point
declared in include/addft_hip.h, and the Python binding covers them all."""

from __future__ import annotations

import ctypes
import re

import pytest

from conftest import ROOT


def _declared():
    text = (ROOT / "include" / "addft_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(addft_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    from addft.kernels import _lib
    path = _lib.lib_path()
    assert path.exists(), f"build the library first: {path}"
    so = ctypes.CDLL(str(path))
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_binding_covers_header():
    from addft.kernels import _lib
    assert set(_declared()) <= set(_lib.exported_symbols())


def test_error_path_without_device():
    """Argument validation happens before any device call."""
    from addft.kernels import _lib
    L = _lib.lib()
    assert L.addft_version() >= 1
    rc = L.addft_pack_scatter(7, 1, 1, 1, 4, 8, 1, None)
    assert rc == -1 and b"dtype" in L.addft_last_error()
    rc = L.addft_pack_scatter(0, None, None, 1, 4, 8, None, None)
    assert rc == -1


def test_autograd_wrappers_import_without_gpu():
    """The autograd wrappers load on a host without a GPU; each defines its
    own backward."""
    import torch
    from addft.kernels import autograd as AG
    assert len(AG.FUNCTIONS) == 7
    for fn in AG.FUNCTIONS:
        assert issubclass(fn, torch.autograd.Function)
        assert fn.backward is not torch.autograd.Function.backward


@pytest.mark.parametrize("name", ["ADDFT_NATIVE_PACK", "ADDFT_NATIVE_XC"])
def test_native_switch_off_is_refused(monkeypatch, name):
    """The reference's "0" switches select a JAX fallback; this build has
    none, so they raise instead of running the HIP path silently."""
    from addft.kernels import _lib
    monkeypatch.setattr(_lib, "_LIB", None)          # the switches are read when the library loads
    monkeypatch.setenv(name, "0")
    with pytest.raises(_lib.NativeUnavailable, match=name):
        _lib.lib()
    monkeypatch.setenv(name, "1")
    monkeypatch.setenv("ADDFT_NATIVE_EIGH", "0")      # accepted: no native dense eigh on this path
    _lib.check_switches()


def test_invalid_arguments_are_reported_without_a_launch():
    """Shape and argument errors come back as ADDFT_EINVAL with a message
    (the reference's handlers return ffi::Error::InvalidArgument,
    pack.cu:50-51, eigh.cu:66-68); the checks run before any HIP call, so
    they hold on a host without a GPU."""
    from addft.kernels import _lib
    L = _lib.lib()
    EINVAL = -1

    def err(rc):
        assert rc == EINVAL
        return L.addft_last_error().decode()

    assert "dtype" in err(L.addft_cheb_combine(7, 1, 10, None, 1.0, 0.0, 0.0, 1, 1, None, 1, None))
    assert "nblk" in err(L.addft_cheb_combine(_lib.C64, 70000, 10, None, 1.0, 0.0, 0.0, 1, 1, None, 1, None))
    assert "null" in err(L.addft_cheb_combine(_lib.C64, 1, 10, None, 1.0, 0.0, 0.0, None, 1, None, 1, None))
    assert "even" in err(L.addft_split_bf16x3(1, 4, 3, 12, 1, None))
    assert "ld" in err(L.addft_split_bf16x3(1, 4, 4, 10, 1, None))
    assert "null" in err(L.addft_pack_scatter(_lib.C64, None, None, 1, 4, 8, None, None))
    assert "size" in err(L.addft_pack_scatter(_lib.C64, 1, 1, 1, 9, 8, 1, None))
    # empty work is a no-op, not an error
    assert L.addft_cheb_combine(_lib.C64, 0, 10, None, 1.0, 0.0, 0.0, None, None, None, None, None) == 0
