"""Build the gfx950 hydro library ``camr_amd/libcamr_hydro.so`` with hipcc.

Every translation unit is compiled as HIP for gfx950 with
``-ffp-contract=off`` (no FMA contraction: the FP64 results must match the
reference CPU build bit for bit) and with packed-FP32 instructions disabled.
Objects go to ``camr_amd/_build``; the shared library is written in-tree so
that it travels with the repository snapshot.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcamr_hydro.so")
# Extra compiler flags (e.g. -D switches) for variant builds; the product
# library is always built without them.
EXTRA = os.environ.get("CAMR_EXTRA_FLAGS", "").split()
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CAMR_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
    "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def _compile(src: str, build_dir: str = BUILD, extra=()) -> tuple[str, str, int]:
    obj = os.path.join(build_dir, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CFLAGS, *extra, "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    return obj, (p.stdout + p.stderr), p.returncode


def build(verbose: bool = False, jobs: int | None = None, out: str = LIB,
          extra=None) -> str:
    extra = EXTRA if extra is None else list(extra)
    build_dir = BUILD if out == LIB else BUILD + "_" + os.path.basename(out).replace(".", "_")
    os.makedirs(build_dir, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    objs, failed = [], []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, log, rc in ex.map(lambda s: _compile(s, build_dir, extra), srcs):
            objs.append(obj)
            if rc != 0:
                failed.append(log)
            elif verbose and log.strip():
                print(log, file=sys.stderr)
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n".join(failed))
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if a != "-v"]
    print(build(verbose="-v" in sys.argv, out=args[0] if args else LIB))
