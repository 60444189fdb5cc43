"""Builds libavbd_hip.so (the HIP kernels for gfx950 and the C ABI of include/avbd.h) in place, with hipcc.

Every HIP source is compiled with -ffp-contract=off (no fused multiply-add, so a step rounds in the reference's order)
and without packed FP32 instructions; the host side is C++17.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libavbd_hip.so")
SOURCES = ["collision.hip", "hploc.hip", "topology.hip", "solve.hip", "scan.hip", "edits.hip", "restitution.hip", "sleeping.hip", "queries.hip", "events.hip", "engine.hip", "capi.cpp", "server.cpp", "hull_cook.cpp"]
ARCH = os.environ.get("AVBD_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    return path


def flags():
    return [
        "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
        f"--offload-arch={ARCH}",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-warning-option",
    ]


def build(verbose: bool = False) -> str:
    out_dir = os.path.join(HERE, "_objs")
    os.makedirs(out_dir, exist_ok=True)
    cc = hipcc()

    def compile_one(src):
        obj = os.path.join(out_dir, src.rsplit(".", 1)[0] + ".o")
        cmd = [cc, *flags(), "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [cc, "-O3", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-Wall", "-c",
                   os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compiling {src} failed:\n{r.stdout}\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"linking libavbd_hip.so failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_addon()
    return LIB


NODE_HEADERS = "/usr/include/node"
ADDON = os.path.join(HERE, "js", "avbd_napi.node")


def build_addon() -> str:
    """The Node N-API addon (js/avbd_napi.cpp) over libavbd_hip.so, when Node's headers are installed."""
    if not os.path.exists(os.path.join(NODE_HEADERS, "node_api.h")):
        return ""
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", f"-I{NODE_HEADERS}",
           os.path.join(HERE, "js", "avbd_napi.cpp"), f"-L{HERE}", "-lavbd_hip", "-Wl,-rpath,$ORIGIN/..",
           "-o", ADDON + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building the N-API addon failed:\n{r.stdout}\n{r.stderr}")
    os.replace(ADDON + ".tmp", ADDON)
    return ADDON


if __name__ == "__main__":
    print(build(verbose=True))
