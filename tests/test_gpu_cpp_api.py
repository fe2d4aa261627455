"""The C++ API (include/dmap_amd.hpp) end to end: the reference's validation
invariants run as a C++ program against the library on the GPU
(tests/cpp/validate_traversal_gpu.cpp, built by dmap_amd/build.py)."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "cpp" / "_build" / "validate_traversal_gpu"


def test_cpp_validate_traversal(torch_gpu):
    if not EXE.exists():
        pytest.fail(f"{EXE} is not built: run python -m dmap_amd.build")
    r = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=240)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "validate_traversal_gpu: PASS" in r.stdout
