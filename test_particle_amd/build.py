"""Build libtpgpu.so (HIP, gfx950) in-tree with hipcc.

    python -m test_particle_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtpgpu.so")
OUT_H5 = os.path.join(HERE, "libtph5.so")
# serial HDF5 (headers and library) for libtph5.so
H5_ROOT = os.environ.get("HDF5_ROOT", "/opt/conda")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: FP64 kernels must perform the same IEEE operations as
# the CPU oracle (no fused multiply-adds other than the explicit fma()).
# -packed-fp32-ops: keep the compiler from emitting packed FP32 math.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "-I", CSRC]
SOURCES = ["tp_gpu.hip", "tp_bs.hip", "tp_mlp.hip", "tp_fieldline.hip", "tp_peak.hip"]
# host-side C (the reference's language), compiled by gcc and linked in
C_SOURCES = ["tp_host.c"]
CC = os.environ.get("CC", "gcc")
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "include")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]}")


def build(verbose=False):
    objs = []
    for c in C_SOURCES:
        o = os.path.join(CSRC, os.path.splitext(c)[0] + ".o")
        _run([CC, *C_FLAGS, "-c", os.path.join(CSRC, c), "-o", o], verbose)
        objs.append(o)
    for h in SOURCES:
        o = os.path.join(CSRC, os.path.splitext(h)[0] + ".o")
        _run([HIPCC, *FLAGS, "-c", os.path.join(CSRC, h), "-o", o], verbose)
        objs.append(o)
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp", *objs],
         verbose)
    os.replace(OUT + ".tmp", OUT)
    build_h5(verbose)
    return OUT


def build_h5(verbose=False):
    """libtph5.so: the HDF5 PIC field input (include/tp_h5.h), linked to the
    serial HDF5 library and to libtpgpu.so beside it."""
    _run([CC, *C_FLAGS, "-I", os.path.join(H5_ROOT, "include"), "-shared", "-o",
          OUT_H5 + ".tmp", os.path.join(CSRC, "tp_host_h5.c"), "-L", HERE, "-ltpgpu",
          "-L", os.path.join(H5_ROOT, "lib"), "-lhdf5",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(H5_ROOT, "lib")], verbose)
    os.replace(OUT_H5 + ".tmp", OUT_H5)
    return OUT_H5


if __name__ == "__main__":
    print(build(verbose=True))
