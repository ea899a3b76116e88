"""Build ``libjrhip.so`` (gfx950) in-tree with hipcc.

``python -m jaxrens_amd.build`` compiles the library next to this file.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
SRCS = [PKG / "csrc" / "jr_lib.hip", PKG / "csrc" / "jr_learned.hip"]
OUT = PKG / "libjrhip.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # canonical FP64 evaluation order: only explicit fma() calls fuse
    "-ffp-contract=off",
    # no packed-FP32 instructions from the compiler
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise FileNotFoundError("hipcc not found")


def _messages(res) -> list[str]:
    return [ln for ln in (res.stdout + res.stderr).splitlines()
            if "packed-fp32-ops" not in ln]


def build(verbose: bool = False) -> Path:
    """Compile every translation unit (in parallel) and link libjrhip.so."""
    objdir = PKG / "csrc" / "build"
    objdir.mkdir(exist_ok=True)
    cc = hipcc()
    procs = []
    for src in SRCS:
        obj = objdir / (src.stem + ".o")
        cmd = [cc, *HIPCC_FLAGS, "-c", "-o", str(obj), str(src)]
        procs.append((src, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                 stderr=subprocess.PIPE, text=True)))
    objs, msgs, failed = [], [], []
    for src, obj, p in procs:
        out, err = p.communicate()
        lines = [ln for ln in (out + err).splitlines() if "packed-fp32-ops" not in ln]
        msgs += lines
        if p.returncode != 0:
            failed.append(src.name)
        objs.append(str(obj))
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}:\n" + "\n".join(msgs))
    tmp = OUT.with_suffix(".so.tmp")
    res = subprocess.run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o",
                          str(tmp), *objs], capture_output=True, text=True)
    msgs += _messages(res)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs), file=sys.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
