"""Builds libvipe_hip.so (gfx950) in-tree: vipe_amd/_lib/libvipe_hip.so.

    python -m vipe_amd.build [-j N] [-v]

Each ``csrc/*.hip`` is compiled to an object with hipcc, in parallel, then
linked into one shared library.  Packed FP32 instruction selection is turned
off for device code (``-packed-fp32-ops``).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT_DIR = HERE / "_lib"
OBJ_DIR = HERE.parent / "_build" / "obj"
LIB = OUT_DIR / "libvipe_hip.so"
ARCH = os.environ.get("VIPE_HIP_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-result",
    "-Wno-unknown-pragmas",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{HERE.parent / 'include'}",
]


def sources():
    return sorted(CSRC.glob("*.hip"))


def _compile(src: Path, verbose: bool):
    obj = OBJ_DIR / (src.stem + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = (r.stdout + r.stderr).strip()
    # hipcc's host pass does not know the packed-fp32 feature; drop that warning only.
    lines = [ln for ln in msg.splitlines() if "packed-fp32-ops" not in ln and "is not a recognized feature" not in ln]
    return src.name, r.returncode, "\n".join(lines), obj


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    OBJ_DIR.mkdir(parents=True, exist_ok=True)
    OUT_DIR.mkdir(parents=True, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    objs = []
    errors = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for name, rc, msg, obj in ex.map(lambda s: _compile(s, verbose), srcs):
            if rc != 0:
                errors.append(f"--- {name} ---\n{msg}")
            elif msg and verbose:
                print(msg)
            objs.append(obj)
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    print(build(a.j, a.v))


if __name__ == "__main__":
    sys.exit(main())
