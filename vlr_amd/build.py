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


# sources compiled once per unit (-DVLRA_UNIT=1..n); each compilation emits one part of the
# file's kernels, so the parts compile in parallel
UNITS = {"kernels.hip": 9}


def _compile(src: Path, obj: Path, debug: bool, defines=()):
    cmd = [hipcc(), "-x", "hip", "-c", str(src), "-o", str(obj)] + _flags(debug) + list(defines)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return src.name, warn


def _deps(src: Path, seen=None) -> set:
    """The file and every quoted include it reaches (csrc/ and include/)."""
    seen = set() if seen is None else seen
    if src in seen or not src.exists():
        return seen
    seen.add(src)
    for line in src.read_text(errors="replace").splitlines():
        line = line.strip()
        if line.startswith("#include \""):
            name = line.split('"')[1]
            for base in (src.parent, PKG.parent / "include"):
                if (base / name).exists():
                    _deps(base / name, seen)
                    break
    return seen


def _stale(src: Path, obj: Path) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return any(t < d.stat().st_mtime for d in _deps(src))


def build(debug: bool = False, jobs: int = 8, verbose: bool = False, incremental: bool = False) -> Path:
    """Compiles every source (incremental=True: only those older than their object or one of
    the headers they include) and links the library."""
    BUILD_DIR.mkdir(exist_ok=True)
    sources = sorted(list(SRC.glob("*.cpp")) + list(SRC.glob("*.hip")))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = []
        for s in sources:
            units = UNITS.get(s.name)
            for u in (range(1, units + 1) if units else [None]):
                o = BUILD_DIR / (s.name + (f".u{u}" if u else "") + ".o")
                objs.append(o)
                if incremental and not debug and not _stale(s, o):
                    continue
                futs.append(ex.submit(_compile, s, o, debug, [f"-DVLRA_UNIT={u}"] if u else []))
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
    build_examples()
    return OUT


EXAMPLES = PKG / "examples"


def build_examples() -> list:
    """Compiles the C and C++ callers in vlr_amd/examples (plain gcc / g++ against the public
    headers, linked to libvlr_amd.so) into vlr_amd/examples/bin."""
    out_dir = EXAMPLES / "bin"
    out_dir.mkdir(exist_ok=True)
    inc = ["-D__HIP_PLATFORM_AMD__", f"-I{PKG.parent / 'include'}", "-I/opt/rocm/include"]
    link = [f"-L{PKG}", "-lvlr_amd", "-L/opt/rocm/lib", "-lamdhip64", "-lm",
            "-Wl,-rpath,$ORIGIN/../..", "-Wl,-rpath,/opt/rocm/lib"]
    built = []
    for src, compiler, std in (("cornell.c", "gcc", "-std=c11"), ("cornell.cpp", "g++", "-std=c++17")):
        exe = out_dir / ("cornell_cpp" if src.endswith(".cpp") else "cornell")
        cmd = [compiler, std, "-O2", "-Wall", *inc, str(EXAMPLES / src), "-o", str(exe), *link]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"example build failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        built.append(exe)
    return built


if __name__ == "__main__":
    print(build(debug="--debug" in sys.argv, verbose=True, incremental="--incremental" in sys.argv))
