"""Build libgacelle_hip.so in-tree with hipcc for gfx950.

Every translation unit under csrc/ (the C ABI in gacelle.hip and one
ops_*.hip per register-tiled model variant) is compiled in parallel to an
object under _build/, then linked into gacelle_amd/libgacelle_hip.so.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ_DIR = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libgacelle_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # no packed-FP32 instructions (v_pk_*_f32) in any HIP code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no SLP vectorisation: it forms <2 x float> phis across uniform
    # switches that this compiler lowers incorrectly (a lane copy is lost)
    "-fno-slp-vectorize",
]


# per translation unit: the dictionary search keeps its FP64 MFMA
# accumulators in VGPRs, so the epilogue reads them without
# v_accvgpr_read copies
EXTRA_FLAGS = {
    "dict.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
}


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def _run(cmd: list[str]) -> tuple[int, str]:
    res = subprocess.run(cmd, capture_output=True, text=True)
    return res.returncode, res.stdout + res.stderr


def _stale(src: str, obj: str) -> bool:
    """The object is missing or older than its source or any shared header."""
    if not os.path.exists(obj):
        return True
    deps = [src, *glob.glob(os.path.join(CSRC, "*.hpp")),
            os.path.join(HERE, "..", "include", "gacelle.h"), os.path.abspath(__file__)]
    return os.path.getmtime(obj) < max(os.path.getmtime(d) for d in deps)


def build_library(verbose: bool = False, jobs: int | None = None, incremental: bool = False) -> str:
    """Compile every translation unit and link the library.  incremental
    recompiles only the objects older than their sources or the headers."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ_DIR, exist_ok=True)
    all_srcs = sources()
    all_objs = [os.path.join(OBJ_DIR, os.path.basename(s)[:-4] + ".o") for s in all_srcs]
    pairs = [(s, o) for s, o in zip(all_srcs, all_objs) if not incremental or _stale(s, o)]
    srcs = [s for s, _ in pairs]
    objs = [o for _, o in pairs]
    jobs = jobs or min(max(len(srcs), 1), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda so: (so[0], *_run([hipcc, *HIPCC_FLAGS,
                                                                       *EXTRA_FLAGS.get(os.path.basename(so[0]), []),
                                                                       "-c", "-o", so[1], so[0]])),
                              zip(srcs, objs)))
    failed = [(s, log) for s, rc, log in results if rc != 0]
    if failed:
        for s, log in failed:
            sys.stderr.write(f"--- {s}\n{log}\n")
        raise RuntimeError(f"hipcc failed on {len(failed)} translation unit(s)")
    rc, log = _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp", *all_objs])
    if rc != 0:
        sys.stderr.write(log)
        raise RuntimeError("linking libgacelle_hip.so failed")
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print(f"built {OUT} from {len(srcs)} translation units")
    return OUT


if __name__ == "__main__":
    build_library(verbose=True, incremental="--incremental" in sys.argv)
