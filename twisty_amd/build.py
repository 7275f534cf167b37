"""Builds libtwisty_hip.so (HIP kernels + C ABI) in-tree for gfx950.

    python -m twisty_amd.build           # build
    python -m twisty_amd.build --asm     # also dump gfx950 assembly to _build/asm

The library lands in twisty_amd/lib/ so that it travels with the source tree.
"""
import argparse
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
BUILD = os.path.join(HERE, "_build")
ROOT = os.path.dirname(HERE)

ARCH = "gfx950"
HIP_SOURCES = ["forward.hip", "histogram.hip", "adjoint.hip", "batched.hip", "generate.hip", "capi.hip", "batch_capi.hip"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 (v_pk_*_f32) stays off: the compiler must not form it.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # No implicit FMA contraction: HIP's __fmul_rn / __fadd_rn are plain
    # operators, so only this keeps the reference's separately rounded float
    # steps (the lookup's lerp, the compensated sums); every fused operation
    # in the kernels is an explicit fmaf.
    "-ffp-contract=off",
    f"-I{os.path.join(ROOT, 'include')}",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(asm=False, verbose=True):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    objs = [os.path.join(BUILD, os.path.splitext(src)[0] + ".o") for src in HIP_SOURCES]
    cmds = [[cc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj] for src, obj in zip(HIP_SOURCES, objs)]
    if asm:
        adir = os.path.join(BUILD, "asm")
        os.makedirs(adir, exist_ok=True)
        cmds += [[cc, *HIPCC_FLAGS, "--cuda-device-only", "-S", os.path.join(CSRC, src), "-o",
                  os.path.join(adir, os.path.splitext(src)[0] + ".s")] for src in HIP_SOURCES]
    # the translation units are independent: compile them side by side
    workers = max(1, min(len(cmds), os.cpu_count() or 1, 8))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
        for f in [pool.submit(run, c) for c in cmds]:
            f.result()
    out = os.path.join(LIBDIR, "libtwisty_hip.so")
    tmp = out + ".tmp"
    run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    build_host()
    calib = os.path.join(LIBDIR, "twisty_copy_calib")
    run([cc, *HIPCC_FLAGS, "-o", calib + ".tmp", os.path.join(CSRC, "copy_calib.hip")])
    os.replace(calib + ".tmp", calib)
    rate = os.path.join(LIBDIR, "twisty_valu_rate")
    run([cc, *HIPCC_FLAGS, "-o", rate + ".tmp", os.path.join(CSRC, "valu_rate.hip")])
    os.replace(rate + ".tmp", rate)
    return out


HOST_SOURCES = ["weighting.cpp", "geometry.cpp", "json.cpp", "runner.cpp", "export.cpp", "batch_runner.cpp",
                "path_generation.cpp", "gradients.cpp",
                "host_capi.cpp"]
DRIVERS = {"twisty_full_experiment": "full_experiment_main.cpp", "twisty_beam_spread": "beam_spread_main.cpp",
           "twisty_ring_benchmark": "ring_benchmark_main.cpp"}
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
              f"-I{os.path.join(ROOT, 'include')}"]


def build_host():
    """libtwisty_host.so (C++ host layer) and the driver executables."""
    cxx = os.environ.get("CXX", "g++")
    hdir = os.path.join(CSRC, "host")
    srcs = [os.path.join(hdir, s) for s in HOST_SOURCES]
    link = [f"-L{LIBDIR}", "-ltwisty_hip", "-Wl,-rpath,$ORIGIN", "-lpthread"]
    out = os.path.join(LIBDIR, "libtwisty_host.so")
    run([cxx, *HOST_FLAGS, "-shared", "-o", out + ".tmp", *srcs, *link])
    os.replace(out + ".tmp", out)
    for name, main_src in DRIVERS.items():
        exe = os.path.join(LIBDIR, name)
        run([cxx, *HOST_FLAGS, "-o", exe + ".tmp", os.path.join(hdir, main_src), "-ltwisty_host", *link])
        os.replace(exe + ".tmp", exe)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--asm", action="store_true")
    a = ap.parse_args(argv)
    print(build(asm=a.asm))


if __name__ == "__main__":
    sys.exit(main())
