"""Build recipe for the doma_amd native library (gfx950).

``python -m doma_amd.build`` compiles every HIP/C++ source under
``doma_amd/csrc`` into ``doma_amd/libdoma_amd.so`` with hipcc, in tree, so the
built library travels with the source tree.  Objects are cached under
``doma_amd/_build`` and rebuilt when a source or header is newer.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
FRONTEND = PKG / "frontend"
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"
LIB = PKG / "libdoma_amd.so"
CLI = PKG / "doma"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DOMA_OFFLOAD_ARCH", "gfx950")

# Packed FP32 VALU ops are disabled for every device pass (they are not used by
# hand either); the host pass warns that it does not know the feature.
COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}", f"-I{CSRC}", f"-I{FRONTEND}",
]


# Per-source flags.  The network's convolution kernels keep their MFMA
# accumulators in ordinary VGPRs: with AGPR accumulators the register allocator
# shuffles them through v_accvgpr moves inside the k-loop.
EXTRA = {
    "unet.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
}


def _sources():
    # The command-line driver (doma_main.cpp) is linked separately.
    srcs = list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")) + list(FRONTEND.glob("*.cpp"))
    return sorted(s for s in srcs if s.name != "doma_main.cpp")


def _headers():
    return sorted(list(CSRC.glob("*.h")) + list(FRONTEND.glob("*.h")) + list(INCLUDE.glob("*.h")))


def _compile(src: Path, force: bool) -> Path:
    obj = BUILD / (src.name + ".o")
    newest_dep = max([src.stat().st_mtime] + [h.stat().st_mtime for h in _headers()])
    if not force and obj.exists() and obj.stat().st_mtime >= newest_dep:
        return obj
    cmd = [HIPCC, *COMMON, *EXTRA.get(src.name, []), "-x", "hip", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed on {src.name}")
    return obj


def build(force: bool = True, jobs: int | None = None) -> Path:
    """Compile and link libdoma_amd.so; returns its path."""
    BUILD.mkdir(exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(8, len(srcs))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("link of libdoma_amd.so failed")
    os.replace(tmp, LIB)
    _link_cli(force)
    return LIB


def _link_cli(force: bool) -> Path:
    """Build the ``doma`` command-line driver against libdoma_amd.so ($ORIGIN rpath)."""
    src = FRONTEND / "doma_main.cpp"
    obj = _compile(src, force)
    tmp = CLI.with_name("doma.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-o", str(tmp), str(obj), f"-L{PKG}", "-ldoma_amd",
           "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("link of the doma driver failed")
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
