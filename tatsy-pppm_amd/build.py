"""Build libpppm_hip.so (gfx950 kernels + C ABI + scene generator) in-tree.

Every HIP translation unit is compiled with hipcc for gfx950 only, with packed
FP32 instructions disabled, and linked into one shared library next to this
file.  `build_all()` also builds the oracle (the CPU checker) under oracle/.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpppm_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include")]

HIP_SOURCES = ["radix.hip", "hash_grid.hip", "bvh.hip", "render.hip", "gather.hip", "proposal.hip", "subsurface.hip",
               "capi.hip", "pass.hip", "sss_host.hip", "proposal_host.hip", "batch.hip", "noalloc.hip",
               "envmap_host.hip", "hachisuka.hip", "dipole_cut.hip"]
CXX_SOURCES = ["scenegen.cc", "pppm_amd.cc", "mesh_io.cc", "image_io.cc"]
FACADE_TEST = os.path.join(ROOT, "tests", "cpp", "facade_main")
# C++ programs over the reference-shaped interface, built beside their sources
CPP_TESTS = ["facade_main", "appmain_shape", "mesh_main", "image_main"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_lib(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or int(os.environ.get("MAX_JOBS", min(8, os.cpu_count() or 4)))
    cmds = []
    objs = []
    for s in HIP_SOURCES:
        src = os.path.join(HERE, "csrc", s)
        obj = os.path.join(BUILD, s + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
        objs.append(obj)
    for s in CXX_SOURCES:
        src = os.path.join(HERE, "host", s)
        obj = os.path.join(BUILD, s + ".o")
        cmds.append(["g++"] + CXX_FLAGS + ["-c", src, "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # the C++ interface driven like the reference's application (test binaries)
    for name in CPP_TESTS:
        exe = os.path.join(ROOT, "tests", "cpp", name)
        _run(["g++"] + CXX_FLAGS + [exe + ".cc", "-o", exe, "-L" + HERE, "-lpppm_hip", "-pthread",
                                    "-Wl,-rpath,$ORIGIN/../../tatsy-pppm_amd"])
    return LIB


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"])
    if os.path.isdir("/root/reference/sources"):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])


def build_all(verbose=False):
    build_oracle()
    return build_lib(verbose=verbose)


if __name__ == "__main__":
    print(build_all(verbose="-v" in sys.argv))
