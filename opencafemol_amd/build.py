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
HIP_SOURCES = ["kernels.hip", "engine.cpp", "capi.cpp"]
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
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp_lib, *objs])
        shutil.move(tmp_lib, os.path.join(LIB_DIR, "libocm.so"))
    if verbose:
        print("built", os.path.join(LIB_DIR, "libocm.so"))


def build_oracle(verbose=False):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if verbose:
        print("built", os.path.join(ROOT, "oracle", "liboracle.so"))


def main(argv):
    if "--hip" in argv:
        build_hip(True)
    elif "--oracle" in argv:
        build_oracle(True)
    else:
        build_oracle(True)
        build_hip(True)


if __name__ == "__main__":
    main(sys.argv[1:])
