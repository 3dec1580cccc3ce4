"""Build recipes of the HIP libraries (gfx950).

``build_core()`` compiles the runtime and the projection/spike kernels into
``annarchy_amd/lib/libannarchy_hip.so``; ``build_model()`` compiles one
generated neuron-model or synapse-model source into its own shared object.  Both call hipcc
directly (no CMake): every kernel is written for CDNA4 wave64.

Flags common to every device build:
  * ``-ffp-contract=off`` on the model libraries so the FP64 neuron updates
    round exactly like the CPU reference loop (no fused multiply-add);
  * ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 math;
  * ``-munsafe-fp-atomics``: hardware FP32/FP64 atomic add for the spike scatter.
"""

from __future__ import annotations

import hashlib
import os
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB_DIR = PKG / "lib"
INCLUDE = ROOT / "include"
GEN_DIR = PKG / "_generated"
CORE_LIB = LIB_DIR / "libannarchy_hip.so"

OFFLOAD_ARCH = os.environ.get("AHIP_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CORE_SOURCES = [
    "runtime.cpp",
    "containers.cpp",
    "batch.cpp",
    "k_rate_psp.hip",
    "k_spike.hip",
    "k_misc.hip",
    "k_batched.hip",
    "k_mlp.hip",
    "k_delay.hip",
    "k_connect.hip",
]

COMMON_FLAGS = [
    f"--offload-arch={OFFLOAD_ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
    "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(
            "command failed: " + " ".join(cmd) + "\n" + res.stdout[-4000:] + res.stderr[-8000:]
        )


def _objects_dir() -> Path:
    d = PKG / "_build"
    d.mkdir(parents=True, exist_ok=True)
    return d


def build_core(jobs: int = 8) -> Path:
    """Compile every core source (in parallel) and link libannarchy_hip.so."""
    from concurrent.futures import ThreadPoolExecutor

    LIB_DIR.mkdir(parents=True, exist_ok=True)
    objdir = _objects_dir()
    srcs = [CSRC / s for s in CORE_SOURCES if (CSRC / s).exists()]

    def compile_one(src: Path) -> Path:
        obj = objdir / (src.name + ".o")
        lang = ["-x", "hip"]
        _run([HIPCC, *COMMON_FLAGS, *lang, "-c", str(src), "-o", str(obj)])
        return obj

    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = CORE_LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={OFFLOAD_ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs), "-ldl"])
    os.replace(tmp, CORE_LIB)
    return CORE_LIB


# The library name hashes the generated source together with the headers it
# is compiled against (the plug-in ABI and the device helpers), so a library
# built against other headers is never picked up under the same name.
_ABI_HEADERS = (INCLUDE / "annarchy_hip_model.h", INCLUDE / "annarchy_hip_synapse.h", CSRC / "ahip_device.hpp")


def _headers_digest() -> bytes:
    h = hashlib.sha1()
    for f in _ABI_HEADERS:
        h.update(f.read_bytes())
    return h.digest()


def model_lib_path(name: str, source: str, kind: str = "model") -> Path:
    h = hashlib.sha1(source.encode() + _headers_digest()).hexdigest()[:12]
    return LIB_DIR / f"libahip_{kind}_{name}_{h}.so"


def prune_model_libs(keep) -> int:
    """Delete model and synapse libraries (and their generated sources) not
    in `keep`; returns how many libraries were removed."""
    keep = {Path(k).name for k in keep}
    n = 0
    for f in LIB_DIR.glob("libahip_*.so"):
        if f.name not in keep:
            f.unlink()
            n += 1
            g = GEN_DIR / (f.stem + ".hip")
            if g.exists():
                g.unlink()
    for g in GEN_DIR.glob("libahip_*.hip"):
        if g.stem + ".so" not in keep:
            g.unlink()
    return n


def build_model(name: str, source: str, force: bool = False, kind: str = "model") -> Path:
    """Write the generated source and compile it into a model library
    (kind "model": a neuron model, "syn": a spiking-synapse model)."""
    GEN_DIR.mkdir(parents=True, exist_ok=True)
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    out = model_lib_path(name, source, kind)
    src = GEN_DIR / (out.stem + ".hip")
    src.write_text(source)
    if out.exists() and not force:
        return out
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, *COMMON_FLAGS, "-ffp-contract=off", "-x", "hip", "-shared", str(src), "-o", str(tmp)])
    os.replace(tmp, out)
    return out
