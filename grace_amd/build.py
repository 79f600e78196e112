"""Build libgrmhd_mi355x.so (gfx950) in-tree with hipcc.

    python -m grace_amd.build            # build the HIP library
    python -m grace_amd.build --oracle   # also build the CPU oracle (test infrastructure)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgrmhd_mi355x.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GRMHD_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: hipcc must not generate v_pk_{add,mul,fma}_f32 (some of
# their forms have given silently wrong results on gfx950); the host pass warns
# that it does not know the feature, which is harmless.
HIP_FLAGS = [
    "-std=c++17", "-O3", f"--offload-arch={ARCH}", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

SOURCES = ["grmhd.hip", "mlp.hip", "setup.cpp"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if out:
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}")


def build_lib():
    srcs = []
    for s in SOURCES:
        p = os.path.join(CSRC, s)
        if os.path.exists(p):
            srcs += ["-x", "hip", p]
    tmp = LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-o", tmp] + srcs)
    os.replace(tmp, LIB)
    return LIB


IO_LIB = os.path.join(HERE, "libgrmhd_io.so")
HDF5_PREFIX = os.environ.get("GRMHD_HDF5_PREFIX", "/opt/conda")


def build_io():
    """libgrmhd_io.so: HDF5 state dumps (host code, linked against HDF5)."""
    inc = os.path.join(HDF5_PREFIX, "include")
    libdir = os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"HDF5 not found under {HDF5_PREFIX}: libgrmhd_io.so not built", flush=True)
        return None
    tmp = IO_LIB + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I" + inc, "-o", tmp,
          os.path.join(CSRC, "state_io.cpp"), "-L" + libdir, "-lhdf5", "-Wl,-rpath," + libdir])
    os.replace(tmp, IO_LIB)
    return IO_LIB


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-j4"])


HOST_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_host_api.cpp")
HOST_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_host_api")


def build_host_test():
    """The C++ test of grace_amd/host/grace_mi355x.hh (links the library and
    the CPU oracle; run on the GPU by tests/test_gpu_cpp_host.py).  Needs the
    oracle built first."""
    os.makedirs(os.path.dirname(HOST_TEST_BIN), exist_ok=True)
    tmp = HOST_TEST_BIN + ".tmp"
    _run([HIPCC, "-std=c++17", "-O2", f"--offload-arch={ARCH}", "-o", tmp, "-x", "hip", HOST_TEST_SRC,
          "-L" + HERE, "-lgrmhd_mi355x", "-L" + os.path.join(ROOT, "oracle", "_build"), "-loracle",
          "-Wl,-rpath,$ORIGIN/../../../grace_amd", "-Wl,-rpath,$ORIGIN/../../../oracle/_build"])
    os.replace(tmp, HOST_TEST_BIN)
    return HOST_TEST_BIN


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_lib()
    build_io()
    if "--oracle" in argv:
        build_oracle()
        build_host_test()


if __name__ == "__main__":
    main()
