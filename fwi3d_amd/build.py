"""Build script for the native parts of fwi3d_amd.

    python -m fwi3d_amd.build                  # HIP library + host library + drivers
    python -m fwi3d_amd.build --incremental    # recompile only changed HIP sources

Outputs (in-tree, git-ignored):
    fwi3d_amd/libfwi3d_hip.so     hand-written gfx950 kernels + C ABI (include/fwi3d.h)
    fwi3d_amd/libfwi3d_host.so    C++ Modeling / Migration / Inversion classes
    bin/modeling_main, bin/migration_main, bin/inversion_main   drivers
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
HIP_LIB = os.path.join(PKG, "libfwi3d_hip.so")
HOST_LIB = os.path.join(PKG, "libfwi3d_host.so")
BIN = os.path.join(ROOT, "bin")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# -ffp-contract=off: one rounding per written operation (the arithmetic
# contract shared with the CPU oracle).  Packed-FP32 instructions are
# disabled for all HIP code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
HIP_SOURCES = ["abi.hip", "kernels_step.hip", "step_inst_fwd.hip", "step_inst_rtm.hip", "step_inst_fwi.hip",
               "step_inst_batch.hip", "step_inst_prefetch.hip", "step_inst_big.hip", "step_inst_two.hip", "step_inst_pair.hip",
               "kernels_aux.hip", "unet_kernels.hip", "unet_net.hip", "noalloc.hip"]

CXX = shutil.which("g++") or "g++"
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wextra", "-ffp-contract=off",
              "-march=x86-64-v3"]
HOST_SOURCES = ["admin.cpp", "geometry.cpp", "modeling.cpp", "migration.cpp", "inversion.cpp", "dataset.cpp",
                "capi.cpp"]
DRIVERS = ["modeling_main", "migration_main", "inversion_main", "setup_main", "preprocess_main"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build_hip(incremental=False):
    """Compile the HIP sources (in parallel) and link libfwi3d_hip.so.  With
    incremental=True the objects are kept in fwi3d_amd/_build and a source is
    recompiled only when it, or a header of csrc/ or include/, is newer than
    its object."""
    from concurrent.futures import ThreadPoolExecutor
    odir = os.path.join(PKG, "_build") if incremental else CSRC
    os.makedirs(odir, exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    headers += [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    newest_header = max(os.path.getmtime(h) for h in headers)
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(odir, src.replace(".hip", ".o"))
        objs.append(obj)
        if incremental and os.path.exists(obj) and os.path.getmtime(obj) > max(os.path.getmtime(path),
                                                                                newest_header):
            continue
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj])
    with ThreadPoolExecutor(max_workers=len(jobs) or 1) as ex:
        list(ex.map(_run, jobs))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", HIP_LIB + ".tmp"])
    os.replace(HIP_LIB + ".tmp", HIP_LIB)
    if not incremental:
        for o in objs:
            os.remove(o)
    return HIP_LIB


def build_host():
    srcs = [os.path.join(HOST, s) for s in HOST_SOURCES if os.path.exists(os.path.join(HOST, s))]
    if not srcs:
        return None
    _run([CXX] + HOST_FLAGS + ["-shared"] + srcs + ["-o", HOST_LIB + ".tmp",
         "-L" + PKG, "-lfwi3d_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    os.makedirs(BIN, exist_ok=True)
    for d in DRIVERS:
        src = os.path.join(HOST, d + ".cpp")
        if os.path.exists(src):
            _run([CXX] + HOST_FLAGS + [src, "-o", os.path.join(BIN, d), "-L" + PKG, "-lfwi3d_host",
                 "-lfwi3d_hip", "-Wl,-rpath,$ORIGIN/../fwi3d_amd"])
    return HOST_LIB


def build_all(incremental=False):
    build_hip(incremental)
    build_host()


if __name__ == "__main__":
    build_all(incremental="--incremental" in sys.argv[1:])
    print("built", HIP_LIB)
