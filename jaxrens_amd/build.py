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
# every translation unit under csrc/ (the walk kernels are split by configuration
# so that they compile in parallel)
SRCS = sorted((PKG / "csrc").glob("*.hip"))
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
    jobs = max(1, min(len(SRCS), int(os.environ.get("MAX_JOBS", "0") or 0)
                      or os.cpu_count() or 1, 16))
    # largest sources first, so the longest compiles start early
    order = sorted(SRCS, key=lambda p: -p.stat().st_size)
    objs = [str(objdir / (src.stem + ".o")) for src in SRCS]
    msgs, failed, running = [], [], []

    def reap(entry):
        # compiler output goes to a log file per source (no pipe can fill up
        # while another compile is being waited for)
        src, p, log = entry
        p.wait()
        text = log.read_text(errors="replace") if log.exists() else ""
        msgs.extend(ln for ln in text.splitlines() if "packed-fp32-ops" not in ln)
        if p.returncode != 0:
            failed.append(src.name)

    for src in order:
        while len(running) >= jobs:
            done = next((e for e in running if e[1].poll() is not None), None)
            if done is None:
                running[0][1].wait()
                continue
            running.remove(done)
            reap(done)
        obj = objdir / (src.stem + ".o")
        log = objdir / (src.stem + ".log")
        cmd = [cc, *HIPCC_FLAGS, "-c", "-o", str(obj), str(src)]
        with open(log, "w") as fh:
            running.append((src, subprocess.Popen(cmd, stdout=fh, stderr=subprocess.STDOUT),
                            log))
    for e in running:
        reap(e)
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
