"""Build recipe for the gfx950 library and the CLI programs.

    python voss-prime-gaps_v1_amd/build.py          # everything
    python voss-prime-gaps_v1_amd/build.py lib      # libvoss_amd.so only

Outputs (in-tree and git-ignored; the tree carries them to where it runs):
    voss-prime-gaps_v1_amd/lib/libvoss_amd.so   HIP kernels + C ABI
    voss-prime-gaps_v1_amd/bin/voss_v7_live     reference-compatible sieve CLI
    voss-prime-gaps_v1_amd/bin/verify_cpp       reference-compatible verifier CLI
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIB_DIR = os.path.join(PKG, "lib")
BIN_DIR = os.path.join(PKG, "bin")
LIB = os.path.join(LIB_DIR, "libvoss_amd.so")
ARCH = os.environ.get("VOSS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions anywhere (the kernels are integer-only)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]
KERNEL_SOURCES = ["sieve.hip", "mr.hip", "api.hip", "serve.hip", "serve_steps.hip", "serve_kernel.hip"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_lib():
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in KERNEL_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", *srcs, "-o", tmp,
          "-Wl,-rpath,/opt/rocm/lib", "-Wl,--no-undefined"])
    os.replace(tmp, LIB)
    return LIB


PROBE = os.path.join(LIB_DIR, "libvoss_probe.so")


def build_probe():
    """LDS roofline microbenchmarks (bench.py reads the measured peak)."""
    os.makedirs(LIB_DIR, exist_ok=True)
    _run([HIPCC, *HIP_FLAGS, "-shared", os.path.join(CSRC, "probe.hip"), "-o", PROBE,
          "-Wl,-rpath,/opt/rocm/lib"])
    return PROBE


def build_cli():
    """C++ host programs over the C ABI (no device code of their own)."""
    os.makedirs(BIN_DIR, exist_ok=True)
    common = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
              "-L", LIB_DIR, "-Wl,-rpath,$ORIGIN/../lib", "-Wl,-rpath,/opt/rocm/lib"]
    outs = []
    for src, exe in (("voss_v7_live.cpp", "voss_v7_live"), ("verify_cpp.cpp", "verify_cpp")):
        path = os.path.join(HOST, src)
        if not os.path.exists(path):
            continue
        out = os.path.join(BIN_DIR, exe)
        _run(common + [path, "-o", out, "-lvoss_amd"])
        outs.append(out)
    return outs


def build_all():
    build_lib()
    build_probe()
    build_cli()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"lib": build_lib, "probe": build_probe, "cli": build_cli, "all": build_all}[what]()
