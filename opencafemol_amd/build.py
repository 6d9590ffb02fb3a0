"""In-tree build of the HIP library (gfx950) and of the CPU oracle.

    python -m opencafemol_amd.build          # both
    python -m opencafemol_amd.build --hip    # libocm.so only

Outputs: opencafemol_amd/lib/libocm.so and oracle/liboracle.so.
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
HIP_SOURCES = ["kernels.hip", "engine.cpp", "server.cpp", "capi.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed-FP32 instructions are disabled for every device object.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops' is not a recognized" not in l
                    and "warning: unknown" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_hip(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        objs = []
        procs = []
        for src in HIP_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            objs.append(obj)
            cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        for cmd, p in procs:
            out, _ = p.communicate()
            if p.returncode != 0:
                raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
        tmp_lib = os.path.join(tmp, "libocm.so")
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-pthread", "-o", tmp_lib, *objs,
              "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx"])
        shutil.move(tmp_lib, os.path.join(LIB_DIR, "libocm.so"))
    if verbose:
        print("built", os.path.join(LIB_DIR, "libocm.so"))


HOST_SOURCES = ["toml.cpp", "forcefield.cpp", "aicg2_tables.cpp", "read_toml.cpp", "read_genesis.cpp", "simulator.cpp", "host_capi.cpp"]
HOST = os.path.join(HERE, "host")
BIN_DIR = os.path.join(HERE, "bin")
CXX = os.environ.get("CXX", "g++")
ROCM_LIB = "/opt/rocm/lib"


def build_host(verbose=False):
    """Host front end (C++17, g++): libocm_host.so over libocm.so, and the
    ocm_run command-line driver."""
    os.makedirs(BIN_DIR, exist_ok=True)
    lib = os.path.join(LIB_DIR, "libocm_host.so")
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-Wno-unused-parameter",
          *[os.path.join(HOST, f) for f in HOST_SOURCES], "-L" + LIB_DIR, "-locm", "-L" + ROCM_LIB,
          "-Wl,-rpath-link," + ROCM_LIB, "-Wl,-rpath,$ORIGIN", "-o", lib])
    _run([CXX, "-O2", "-std=c++17", os.path.join(HOST, "main.cpp"), "-L" + LIB_DIR, "-locm_host", "-locm",
          "-L" + ROCM_LIB, "-Wl,-rpath-link," + ROCM_LIB, "-Wl,-rpath,$ORIGIN/../lib",
          "-o", os.path.join(BIN_DIR, "ocm_run")])
    if verbose:
        print("built", lib, "and", os.path.join(BIN_DIR, "ocm_run"))


def build_oracle(verbose=False):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if verbose:
        print("built", os.path.join(ROOT, "oracle", "liboracle.so"))


def main(argv):
    if "--hip" in argv:
        build_hip(True)
    elif "--oracle" in argv:
        build_oracle(True)
    elif "--host" in argv:
        build_host(True)
    else:
        build_oracle(True)
        build_hip(True)
        build_host(True)


if __name__ == "__main__":
    main(sys.argv[1:])
