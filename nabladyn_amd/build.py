"""Build of the gfx950 library: every .hip under csrc/ compiled by hipcc, linked into one .so.

Run as ``python -m nabladyn_amd.build`` or through ``__graft_entry__.build()``. It compiles on
a machine without a GPU (hipcc cross-compiles for gfx950). The library lands in
``nabladyn_amd/_lib/libnabladyn_amd.so``, in-tree, where ``nabladyn_amd._native`` loads it.
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent
CSRC = ROOT / "csrc"
OUT = ROOT / "_lib"
LIBRARY = OUT / "libnabladyn_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-fno-gpu-rdc",
    # packed fp32 instructions off (see the header of csrc/common.hpp for why)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall",
    "-Wno-unused-result",
]
# The neighbor list is bit-exact only without contraction; the force kernels keep their
# distance test exact with `#pragma clang fp contract(off)` and may fuse elsewhere.
PER_FILE = {
    "neighbors.hip": ["-ffp-contract=off"],
}
DEFAULT_FP = ["-ffp-contract=fast-honor-pragmas"]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def _compile(source: Path) -> Path:
    obj = OUT / "obj" / (source.stem + ".o")
    obj.parent.mkdir(parents=True, exist_ok=True)
    flags = COMMON + PER_FILE.get(source.name, DEFAULT_FP)
    cmd = [HIPCC, *flags, "-c", str(source), "-o", str(obj)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"hipcc failed on {source.name}:\n{result.stderr}")
    if result.stderr.strip():
        lines = [l for l in result.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            print(f"[{source.name}]\n" + "\n".join(lines), file=sys.stderr)
    return obj


def build(verbose: bool = True) -> Path:
    """Compile every source and link the library; return its path."""
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as pool:
        objs = list(pool.map(_compile, srcs))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"link failed:\n{result.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY} from {len(srcs)} sources")
    return LIBRARY


if __name__ == "__main__":
    build()
