"""Build libvkrt.so (gfx950 HIP kernels + C ABI) in-tree, and the CPU oracle beside it.

Run as ``python -m vierkant_amd.build`` or through ``__graft_entry__.build()``. Every HIP source is
compiled with -ffp-contract=off (the reference arithmetic has no fused multiply-add), IEEE division and
square root (hipcc's defaults), and with packed FP32 instructions disabled.
"""
import concurrent.futures
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libvkrt.so"
ARCH = "gfx950"

HIPCC_FLAGS = [
    "-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    f"-I{ROOT / 'include'}", f"-I{PKG / 'include'}", f"-I{CSRC}",
]


def _hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libvkrt.so")
    return exe


def _sources():
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src):
    obj = BUILD / (src.name + ".o")
    cmd = [_hipcc(), *HIPCC_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass does not know the packed-fp32 target feature and says so; that is harmless
    noise = [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n" + "\n".join(noise))
    return obj, "\n".join(noise)


def build_lib(verbose=False):
    BUILD.mkdir(exist_ok=True)
    srcs = _sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        results = list(ex.map(_compile, srcs))
    for obj, warn in results:
        if verbose and warn:
            print(warn, file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *[str(o) for o, _ in results],
           "-o", str(tmp), "-lpthread", "-lz"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link of libvkrt.so failed:\n" + res.stderr)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    """CPU oracle (test infrastructure) and, when the reference checkout is present, oracle/_ref."""
    res = subprocess.run(["make", "-C", str(ROOT / "oracle")], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stdout + res.stderr)


def build_test_helpers():
    """tests/cpp/librt_math_host.so: the intersection header compiled for the host (CPU tests)"""
    src = ROOT / "tests" / "cpp" / "rt_math_host.cpp"
    if not src.exists():
        return
    out = src.parent / "librt_math_host.so"
    cmd = ["g++", "-O3", "-std=c++20", "-fPIC", "-shared", "-ffp-contract=off", f"-I{PKG / 'include'}",
           str(src), "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("host build of rt_math.hpp failed:\n" + res.stderr)


def build_host_api_test():
    """tests/cpp/test_host_api: the C++ host layer (vierkant/*.hpp) linked against libvkrt.so"""
    src = ROOT / "tests" / "cpp" / "test_host_api.cpp"
    if not src.exists():
        return
    out = src.parent / "test_host_api"
    cmd = ["g++", "-O2", "-std=c++20", "-ffp-contract=off", f"-I{ROOT / 'include'}", f"-I{PKG / 'include'}",
           str(src), "-o", str(out), f"-L{PKG}", "-lvkrt", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../../vierkant_amd"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("build of the C++ host-layer test failed:\n" + res.stderr)


def build_all(verbose=False):
    build_oracle()
    build_test_helpers()
    lib = build_lib(verbose=verbose)
    build_host_api_test()
    return lib


if __name__ == "__main__":
    print(build_all(verbose=True))
