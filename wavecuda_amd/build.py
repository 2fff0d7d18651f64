"""Build libwavehip.so (HIP, gfx950) in place: ``python -m wavecuda_amd.build``.

Every source under ``csrc/`` is compiled by hipcc for gfx950 and linked
into ``wavecuda_amd/libwavehip.so``.  The packed-FP32 feature is switched
off for device code (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libwavehip.so"
OBJ = HERE / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
CFLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def _sources():
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _stale(src: Path, obj: Path) -> bool:
    """The object is missing or older than its source or any header here."""
    if not obj.exists():
        return True
    newest = max([src.stat().st_mtime] + [h.stat().st_mtime for h in CSRC.glob("*.h")] +
                 [(HERE.parent / "include" / "wavehip.h").stat().st_mtime])
    return obj.stat().st_mtime < newest


def _compile(src: Path, incremental: bool = False) -> Path:
    obj = OBJ / (src.name + ".o")
    if incremental and not _stale(src, obj):
        return obj
    cmd = [HIPCC, *CFLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr}")
    return obj


def build(verbose: bool = False, incremental: bool = False) -> Path:
    """Compile every source (with incremental=True only those whose object
    is older than the source or a header) and link the shared library;
    returns its path."""
    OBJ.mkdir(exist_ok=True)
    srcs = _sources()
    jobs = min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
    sys.exit(0)
