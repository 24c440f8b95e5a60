"""The C ABI library loads and exports every entry point include/pci_gpu.h
declares (no device calls)."""
import ctypes
import os
import re

from pci_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "pci_gpu.h")).read()
    return sorted(set(re.findall(r"\b(pci_gpu_\w+)\s*\(", text)))


def test_header_declares_the_abi():
    names = declared()
    assert "pci_gpu_build_h" in names and "pci_gpu_davidson" in names
    assert set(names) == set(_lib.SYMBOLS)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for n in declared():
        assert hasattr(L, n), n
        assert ctypes.cast(getattr(L, n), ctypes.c_void_p).value


def test_host_library_exports():
    from pci_amd import host
    H = host.host()
    for n in ["pconf_load", "pconf_dims", "pconf_params", "pconf_space", "pconf_upload", "pconf_error"]:
        assert hasattr(H, n)


def test_context_creation_fails_loudly_without_a_gpu():
    import pytest
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(_lib.PciError):
        _lib.GpuContext(0)
