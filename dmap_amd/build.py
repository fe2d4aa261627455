"""Build libdmap_amd.so for gfx950 with hipcc, in-tree (dmap_amd/_lib/).

    python -m dmap_amd.build          # or dmap_amd.build.build()

The library is the C ABI of include/dmap_amd.h; Python reaches it through
ctypes (dmap_amd/__init__.py). Nothing here is needed at run time.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT_DIR = HERE / "_lib"
LIB = OUT_DIR / "libdmap_amd.so"
ARCH = os.environ.get("DMAP_AMD_ARCH", "gfx950")

# -packed-fp32-ops: keep the compiler from forming v_pk_*_f32.
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-Wno-unused-command-line-argument", f"-I{CSRC}"]
# The pyramid build keeps the reference's double arithmetic bit for bit.
PER_FILE = {"pyramid.hip": ["-ffp-contract=off"], "traverse.hip": [], "reorder.hip": [], "bvh.hip": [],
            "api.cpp": ["-x", "hip"], "serve.cpp": ["-x", "hip"], "pathtrace.cpp": ["-x", "hip", "-ffp-contract=off"]}


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    OUT_DIR.mkdir(exist_ok=True)
    cc = hipcc()
    objs = []
    jobs = []
    for src, extra in PER_FILE.items():
        obj = OUT_DIR / (src.replace(".", "_") + ".o")
        objs.append(obj)
        jobs.append([cc, *COMMON, *extra, "-c", str(CSRC / src), "-o", str(obj)])

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        return r.stderr

    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        for err in ex.map(run, jobs):
            if verbose and err.strip():
                print(err)
    tmp = LIB.with_suffix(".so.tmp")
    run([cc, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    for o in objs:
        o.unlink()
    return LIB


ROOT = HERE.parent
CPP_TESTS = {"validate_traversal_gpu": ROOT / "tests" / "cpp" / "validate_traversal_gpu.cpp"}


def build_cpp_tests(verbose=False):
    """The C++ programs that exercise the library through include/dmap_amd.hpp
    (tests/cpp/, linked against the built library with g++), into
    tests/cpp/_build/."""
    out_dir = ROOT / "tests" / "cpp" / "_build"
    out_dir.mkdir(parents=True, exist_ok=True)
    gxx = shutil.which("g++") or "g++"
    built = []
    for name, src in CPP_TESTS.items():
        exe = out_dir / name
        cmd = [gxx, "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(src), f"-L{OUT_DIR}", "-ldmap_amd",
               "-Wl,-rpath,$ORIGIN/../../../dmap_amd/_lib", "-o", str(exe)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        built.append(exe)
    return built


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_cpp_tests(verbose="-v" in sys.argv))
