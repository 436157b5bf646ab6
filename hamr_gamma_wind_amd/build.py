"""Builds libhamr_hip.so (gfx950) in-tree with hipcc, and libhamr_host.so (the
C host of the reference's per-block step sequence, gcc, linked to it).

Everything is compiled without FMA contraction (-ffp-contract=off): the FP64
step follows the reference's expression order, and the FP32 fluxes of the mixed
step keep the face-field error of the CT update at FP32 roundoff.  Packed-FP32
VALU instructions are disabled for all device code.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libhamr_hip.so")
HOST_OUT = os.path.join(HERE, "libhamr_host.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HAMR_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-variable",
          "-Wno-unused-function", "-I" + CSRC]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = [
    ("hamr_kernels.hip", True),
    ("hamr_kernels_f32.hip", True),
    ("hamr_adjoint.hip", True),
    ("hamr_rdump.hip", True),
    ("hamr_capi.hip", True),
    ("hamr_setup.cpp", False),
    ("hamr_amr.cpp", False),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src, dev in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + COMMON + (DEVICE if dev else ["-x", "hip", "--offload-arch=" + ARCH])
        cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
        jobs.append((cmd, obj))
    # the units compile concurrently (one hipcc per unit)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1, 8)) as ex:
        outs = list(ex.map(lambda j: _run(j[0]), jobs))
    if verbose:
        for out in outs:
            if out.strip():
                print(out)
    objs = [obj for _, obj in jobs]
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, OUT)
    tmp = HOST_OUT + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-o", tmp, os.path.join(CSRC, "hamr_host.c"), "-L" + HERE,
          "-l:libhamr_hip.so", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
