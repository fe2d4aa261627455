"""The C-ABI library builds, loads and exports every entry point that
include/dmap_amd.h declares (no GPU needed: no compute call is made)."""
import ctypes
import re
from pathlib import Path

import pytest

import dmap_amd

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include" / "dmap_amd.h").read_text()
    return sorted(set(re.findall(r"\b(dmap_[a-z_]+)\s*\(", text)))


def test_header_and_python_agree():
    assert declared() == sorted(dmap_amd.EXPORTS)


def test_library_exports_every_symbol():
    if not dmap_amd.LIB_PATH.exists():
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(dmap_amd.LIB_PATH))
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert not missing
    assert dmap_amd.load_library().dmap_version().decode().startswith("dmap_amd")


def test_no_gpu_raises_instead_of_falling_back():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(dmap_amd.DmapError):
        dmap_amd.Context(0)


def test_product_does_not_touch_oracle():
    """The product package never imports, links or loads the oracle."""
    for p in (ROOT / "dmap_amd").rglob("*"):
        if p.suffix in (".py", ".hip", ".cpp", ".h"):
            text = p.read_text()
            assert "import oracle" not in text and "from oracle" not in text, p
            assert "dmap_oracle" not in text, p


def test_cpp_header_compiles():
    """include/dmap_amd.hpp (the C++ API) compiles on its own with g++."""
    import shutil
    import subprocess
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    gxx = shutil.which("g++")
    assert gxx
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-x", "c++", f"-I{root / 'include'}", "-"],
                       input='#include "dmap_amd.hpp"\nint main() { return 0; }\n', capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
