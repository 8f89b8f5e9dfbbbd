"""The C-ABI libraries load and export every symbol include/*.h declares;
the Python surface rejects bad operands before any GPU work (CPU only)."""
import ctypes
import os
import re

import pytest
import torch

from cbm import _hip_sigs, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:cbm|csr)_\w+)\s*\(", text)))


@pytest.mark.parametrize("header,loader", [("cbm_host.h", _lib.host), ("cbm_hip.h", _lib.hip)])
def test_libraries_export_declared_symbols(header, loader):
    lib = loader()
    names = _declared(header)
    assert names, header
    for name in names:
        assert hasattr(lib, name), f"{name} declared in {header} is not exported"


def test_python_declarations_cover_header():
    assert sorted(_hip_sigs.EXPORTED) == _declared("cbm_hip.h")


def test_desc_layout_matches_c_struct():
    assert ctypes.sizeof(_hip_sigs.CbmDesc) == _lib.hip().cbm_desc_size()


def test_error_strings_are_empty_initially():
    assert isinstance(_lib.host().cbm_host_last_error(), bytes)
    assert isinstance(_lib.hip().cbm_last_error(), bytes)


def test_cpu_tensors_are_rejected_like_the_reference():
    """cusparseCBM / cusparseCSR raise 'tensor is not a CUDA tensor' for host
    tensors (cbm/cbm.cu:101-108, TORCH_CHECK), before touching a GPU."""
    from cbm.cbm_cusparse import cusparseCBM
    from cbm.csr_cusparse import cusparseCSR
    crow = torch.tensor([0, 1, 2], dtype=torch.int32)
    col = torch.tensor([1, 0], dtype=torch.int32)
    a = torch.sparse_csr_tensor(crow, col, torch.ones(2), (2, 2))
    e = torch.zeros(0, dtype=torch.int32)
    with pytest.raises(RuntimeError, match="not a CUDA tensor"):
        cusparseCBM(a, torch.zeros(1, dtype=torch.int32), e, e, torch.empty(0))
    with pytest.raises(RuntimeError, match="not a CUDA tensor"):
        cusparseCSR(a)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_hip", None)
    monkeypatch.setattr(_lib, "HIP_LIB", str(tmp_path / "absent.so"))
    with pytest.raises(_lib.NativeLibraryMissing):
        _lib.hip()
