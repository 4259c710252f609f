"""Build libvlr_amd.so (C-ABI library: host runtime + HIP kernels for gfx950) in-tree.

    python -m vlr_amd.build        # or vlr_amd.build.build()

Every .cpp/.hip file under vlr_amd/csrc is compiled with hipcc for gfx950 and linked into
vlr_amd/libvlr_amd.so next to vlr_amd/data/spectral_tables.bin, which the library loads at
run time.  Device code is compiled without floating-point contraction (closest-hit parity with
the CPU oracle is bit-exact) and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
SRC = PKG / "csrc"
OUT = PKG / "libvlr_amd.so"
BUILD_DIR = PKG / "_build"
ARCH = os.environ.get("VLR_AMD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _flags(debug: bool = False):
    f = ["-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off", "-fvisibility=hidden",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
         "-Wno-unused-variable", "-Wno-unused-but-set-variable", "-Wno-unused-command-line-argument"]
    f += ["-O0", "-g"] if debug else ["-O3"]
    return f


def _compile(src: Path, obj: Path, debug: bool):
    cmd = [hipcc(), "-x", "hip", "-c", str(src), "-o", str(obj)] + _flags(debug)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return src.name, warn


def build(debug: bool = False, jobs: int = 8, verbose: bool = False) -> Path:
    BUILD_DIR.mkdir(exist_ok=True)
    sources = sorted(list(SRC.glob("*.cpp")) + list(SRC.glob("*.hip")))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = []
        for s in sources:
            o = BUILD_DIR / (s.name + ".o")
            objs.append(o)
            futs.append(ex.submit(_compile, s, o, debug))
        for f in futs:
            name, warn = f.result()
            if verbose and warn:
                print(f"[{name}]\n{warn}", file=sys.stderr)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-o", str(tmp)] + [str(o) for o in objs] + [f"--offload-arch={ARCH}", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(debug="--debug" in sys.argv, verbose=True))
