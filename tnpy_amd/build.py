"""Build ``libtnx.so`` in-tree with hipcc for gfx950.

``python -m tnpy_amd.build`` compiles every ``csrc/*.hip`` to an object and
links the shared library next to this file. It always recompiles (the build
check runs it on machines where a stale library may exist).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR / "csrc"
LIB_PATH = PKG_DIR / "libtnx.so"
ARCH = os.environ.get("TNX_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # packed FP32 VALU ops are disabled on this target family
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the tnx library needs ROCm's hipcc to build")


def build(verbose: bool = False, jobs: int = 0) -> Path:
    hipcc = _hipcc()
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    with tempfile.TemporaryDirectory(prefix="tnx_build_") as tmp:
        objs = []

        def compile_one(src: Path) -> Path:
            obj = Path(tmp) / (src.stem + ".o")
            cmd = [hipcc, *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
            proc = subprocess.run(cmd, capture_output=True, text=True)
            if proc.returncode != 0:
                raise RuntimeError(f"hipcc failed on {src.name}:\n{proc.stderr[-4000:]}")
            if verbose:
                print(f"compiled {src.name}", file=sys.stderr)
            return obj

        n = jobs or min(len(sources), max(1, (os.cpu_count() or 2)))
        with ThreadPoolExecutor(max_workers=n) as ex:
            objs = list(ex.map(compile_one, sources))
        tmp_lib = Path(tmp) / "libtnx.so"
        cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp_lib)]
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            raise RuntimeError(f"link failed:\n{proc.stderr[-4000:]}")
        # atomic replace so a concurrently loaded library is never half-written
        staged = LIB_PATH.with_suffix(".so.tmp")
        shutil.copyfile(tmp_lib, staged)
        os.replace(staged, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
