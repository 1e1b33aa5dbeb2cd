"""Build of the native library ``ustb_amd/libustbx.so`` for gfx950.

Each HIP translation unit under ``ustb_amd/csrc`` is compiled with hipcc in
parallel, then linked into one shared library in the package directory (so it
ships with the source tree and is the library the package loads).

Run as ``python -m ustb_amd.build_lib``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libustbx.so")
ARCH = os.environ.get("USTBX_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    # packed FP32 (v_pk_*_f32) stays off: scalar FP32 only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm not installed?)")


def _compile(src: str, obj: str) -> tuple[str, int, str]:
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", src, "-o", obj]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    return src, p.returncode, out


def build(verbose: bool = True, jobs: int | None = None) -> str:
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = jobs or min(len(srcs), max(1, os.cpu_count() or 1), 16)
    with tempfile.TemporaryDirectory(prefix="ustbx_build_") as tmp:
        objs = [os.path.join(tmp, os.path.basename(s)[:-4] + ".o") for s in srcs]
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            results = list(ex.map(lambda so: _compile(*so), zip(srcs, objs)))
        failed = [r for r in results if r[1] != 0]
        for src, rc, out in results:
            if verbose and out.strip():
                print(f"[{os.path.basename(src)}]\n{out}", file=sys.stderr)
        if failed:
            raise RuntimeError("hipcc failed for: " + ", ".join(os.path.basename(f[0]) for f in failed))
        tmp_lib = LIB + ".tmp"
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp_lib, *objs,
               "-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"]
        p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if p.returncode != 0:
            raise RuntimeError("link failed:\n" + p.stdout)
        os.replace(tmp_lib, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
