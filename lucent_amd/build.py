"""Compiles the gfx950 library ``lucent_amd/liblucent_hip.so`` in-tree.

HIP translation units (``*.hip``) go through hipcc for gfx950; host-only
C++ (``*.cpp``) through g++ against the HIP runtime headers.  Every run
recompiles (the build is a few seconds per file).

    python -m lucent_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "liblucent_hip.so"
BUILD_DIR = PKG / "_build"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # Multiply-adds are fused only within one source expression, never across
    # statements: the same path computed by different kernels (per-sample
    # launch, persistent tiles, adjoint replay) then rounds identically.
    "-ffp-contract=on",
    # Packed FP32 (v_pk_*_f32) stays off on this hardware pool.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
# The path-tracing translation units evaluate trace.frag's sin/cos/log with
# the hardware approximations (GLSL leaves their precision to the
# implementation).  All three share the flag, so the same path still rounds
# identically in every kernel.  (Relaxed fp32 division is not used: its
# lowering depends on the surrounding code, which breaks that identity.)
PT_SOURCES = {"pathtrace.hip", "wavefront.hip", "adjoint.hip"}
PT_FLAGS = ["-fgpu-approx-transcendentals"]
CXX_FLAGS = [
    "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}", "-Wall", "-Wno-unknown-pragmas",
]


def _hipcc() -> str:
    p = shutil.which("hipcc") or str(ROCM / "bin" / "hipcc")
    return p


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _compile(src: Path) -> Path:
    obj = BUILD_DIR / (src.name + ".o")
    if src.suffix == ".hip":
        extra = PT_FLAGS if src.name in PT_SOURCES else []
        cmd = [_hipcc(), *HIPCC_FLAGS, *extra, "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build(verbose: bool = True) -> Path:
    BUILD_DIR.mkdir(exist_ok=True)
    sources = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = OUT.with_suffix(".so.tmp")
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          f"-L{ROCM / 'lib'}", "-lamdhip64", f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, OUT)
    build_node_addon()
    if verbose:
        print(f"built {OUT} from {len(sources)} sources (+ node/build/lucent_napi.node)", file=sys.stderr)
    return OUT


NODE_DIR = PKG.parent / "node"
NODE_INCLUDE = Path("/usr/include/node")


def build_node_addon() -> Path | None:
    """The N-API addon (node/build/lucent_napi.node), linked against the
    library with a relative rpath so the tree can move as a whole."""
    if not (NODE_INCLUDE / "node_api.h").exists():
        print("node headers not found: skipping the N-API addon", file=sys.stderr)
        return None
    out_dir = NODE_DIR / "build"
    out_dir.mkdir(exist_ok=True)
    out = out_dir / "lucent_napi.node"
    tmp = out.with_suffix(".node.tmp")
    _run(["gcc", "-O2", "-shared", "-fPIC", "-std=c11", "-Wall", f"-I{NODE_INCLUDE}",
          "-DNODE_GYP_MODULE_NAME=lucent_napi", str(NODE_DIR / "lucent_napi.c"), "-o", str(tmp),
          f"-L{PKG}", "-llucent_hip", "-Wl,-rpath,$ORIGIN/../../lucent_amd"])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build()
