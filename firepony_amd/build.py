"""Build the native libraries in-tree.

    python -m firepony_amd.build            # engine (gfx950) + synth
    python -m firepony_amd.build --oracle   # also the CPU oracle (test infrastructure)

Outputs:
    firepony_amd/_lib/libfirepony_amd.so   HIP engine + C ABI (hipcc, gfx950)
    firepony_amd/_lib/libfpsynth.so        synthetic workload generator (g++)
    firepony_amd/_lib/libfirepony_amd_io.so  FASTA / VCF / SAM / BAM loaders (g++, zlib)
    oracle/_build/libfp_oracle.so          CPU oracle (g++), built by oracle/build.sh
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "firepony_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "_lib")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FP_OFFLOAD_ARCH", "gfx950")

# device sources of the engine; host C++ sources are compiled by hipcc too
ENGINE_SOURCES = [
    "engine.cpp",
    "report.cpp",
    "k_prep.hip",
    "k_baq.hip",
    "k_cov.hip",
    "k_sort.hip",
    "k_table.hip",
    "k_post.hip",
]

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    # bit-exact FP64: no contraction of a*b+c into fma, anywhere
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def build_synth():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libfpsynth.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread",
          os.path.join(CSRC, "synth.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_io():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libfirepony_amd_io.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall", "-I", os.path.join(ROOT, "include"),
          os.path.join(CSRC, "loaders.cpp"), "-lz", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_engine():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libfirepony_amd.so")
    srcs = [os.path.join(CSRC, s) for s in ENGINE_SOURCES]
    cmd = [HIPCC] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-I", CSRC] + srcs + \
          ["-o", out + ".tmp"]
    # hipcc compiles .cpp as HIP when told to
    cmd.insert(1, "-x")
    cmd.insert(2, "hip")
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_oracle():
    script = os.path.join(ROOT, "oracle", "build.sh")
    _run(["bash", script])
    return os.path.join(ROOT, "oracle", "_build", "libfp_oracle.so")


def main(argv):
    build_synth()
    build_io()
    if "--oracle" in argv or "--all" in argv:
        build_oracle()
    if "--no-engine" not in argv:
        build_engine()


if __name__ == "__main__":
    main(sys.argv[1:])
