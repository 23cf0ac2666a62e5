"""Build recipe for the native parts of gsnp_amd.

Compiles every source under gsnp_amd/csrc with hipcc for gfx950 into
gsnp_amd/_build/libgsnp.so (the C-ABI library declared in include/gsnp_abi.h)
and the command-line tools gsnp and gsnp-decompress, in place.  Object files
are cached by content hash so rebuilding an unchanged tree only relinks.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GSNP_OFFLOAD_ARCH", "gfx950")

LIB_SOURCES = ["tables.cpp", "inputs.cpp", "cwriter.cpp", "synth.cpp", "kernels.hip", "dict.hip", "cols.hip", "calib.hip", "engine.hip", "run.hip",
               "adjoint.hip", "noalloc.hip", "cdec.hip"]
TOOLS = {"gsnp": "cli_main.cpp", "gsnp-decompress": "decompress_main.cpp"}

# -ffp-contract=off everywhere: FP64 sums and comparisons must round exactly as
# written.  Packed FP32 is disabled for device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-result",
          f"-I{ROOT / 'include'}", f"-I{CSRC}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-fno-gpu-rdc"]


def _hash(paths: list[Path], flags: list[str]) -> str:
    h = hashlib.sha256()
    for p in paths:
        h.update(p.read_bytes())
    h.update(" ".join(flags).encode())
    return h.hexdigest()[:16]


def _compile(src: Path, obj: Path, flags: list[str], force: bool = False) -> None:
    headers = sorted(CSRC.glob("*.h")) + [ROOT / "include" / "gsnp_abi.h"]
    digest = _hash([src] + headers, flags)
    stamp = obj.with_suffix(".hash")
    if not force and obj.exists() and stamp.exists() and stamp.read_text() == digest:
        return
    cmd = [HIPCC] + flags + ["-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    stamp.write_text(digest)


def build(verbose: bool = True, force: bool = False) -> Path:
    """Compiles the library and tools; unchanged objects are reused unless force."""
    OUT.mkdir(exist_ok=True)
    jobs = []
    for name in LIB_SOURCES:
        src = CSRC / name
        flags = COMMON + (DEVICE + ["-x", "hip"] if name.endswith(".hip") else [])
        jobs.append((src, OUT / (src.stem + ".o"), flags))
    for name in TOOLS.values():
        src = CSRC / name
        jobs.append((src, OUT / (src.stem + ".o"), COMMON))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda j: _compile(*j, force=force), jobs))
    lib = OUT / "libgsnp.so"
    objs = [str(OUT / (Path(n).stem + ".o")) for n in LIB_SOURCES]
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", str(lib)] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    for tool, src in TOOLS.items():
        exe = OUT / tool
        cmd = [HIPCC, "-o", str(exe), str(OUT / (Path(src).stem + ".o")), f"-L{OUT}", "-lgsnp",
               f"-Wl,-rpath,$ORIGIN"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    try:
        build(force="--force" in sys.argv)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
