"""The C-ABI library builds, loads and exports every symbol its header
declares (no device needed)."""
import os
import re

from hsmc_amd import gpu
from tests.conftest import ROOT


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hsmc_\w+)\s*\(", text)))


def test_gpu_library_exports_header():
    L = gpu.load()
    names = _declared("hsmc_gpu.h")
    assert names, "no declarations parsed"
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(gpu.EXPORTS)


def test_strerror_without_device():
    L = gpu.load()
    assert L.hsmc_gpu_strerror(0) == b"ok"
    assert b"cell" in L.hsmc_gpu_strerror(-1)
