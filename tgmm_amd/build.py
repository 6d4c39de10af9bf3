"""Build recipe for the native libraries (hipcc for gfx950, g++ for host).

    python -m tgmm_amd.build        # builds everything in-tree

Outputs (git-ignored, shipped with the tree):
    tgmm_amd/libtgmm_em.so     the EM library (HIP kernels + C-ABI)
    tgmm_amd/libtgmm_synth.so  the synthetic frame generator (host)
    oracle/liboracle.so        the checker (tests and CPU baseline only)
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["tgmm_kernels.hip", "tgmm_runtime.hip", "tgmm_legacy.hip",
               "tgmm_server.hip", "tgmm_local.hip", "tgmm_marshal.hip",
               "tgmm_svknn.hip"]
HIP_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), out))
    return out


def build_em(jobs=None):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("hipcc failed: %s\n%s" % (" ".join(cmd), out))
    so = os.path.join(HERE, "libtgmm_em.so")
    tmp = so + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, so)
    return so


def build_synth():
    so = os.path.join(HERE, "libtgmm_synth.so")
    tmp = so + ".tmp"
    _run(["g++", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-o", tmp,
          os.path.join(CSRC, "synth.cpp")])
    os.replace(tmp, so)
    return so


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B", "all"])
    if os.path.isdir("/root/reference/nmeth.3036-S2/src"):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])


def build_devmath_host():
    """Host build of the kernels' per-Gaussian math, for CPU tests only."""
    src = os.path.join(ROOT, "tests", "native", "devmath_host.cpp")
    if not os.path.exists(src):
        return None
    so = os.path.join(ROOT, "tests", "native", "libdevmath_host.so")
    tmp = so + ".tmp"
    _run([HIPCC, "-O2", "-fPIC", "-shared", "-std=c++17", "-o", tmp, src])
    os.replace(tmp, so)
    return so


def build_compat_frame():
    """The C++ test program that runs a frame through the reference's names
    (include/*_compat.hpp), for tests/test_gpu_compat.py."""
    src = os.path.join(ROOT, "tests", "native", "compat_frame.cpp")
    if not os.path.exists(src):
        return None
    exe = os.path.join(ROOT, "tests", "native", "compat_frame")
    tmp = exe + ".tmp"
    _run(["g++", "-O2", "-std=c++11", "-I", os.path.join(ROOT, "include"), "-o", tmp, src,
          "-L", HERE, "-ltgmm_em", "-Wl,-rpath,$ORIGIN/../../tgmm_amd",
          "-Wl,-rpath-link,/opt/rocm/lib"])
    os.replace(tmp, exe)
    return exe


def build_all():
    build_synth()
    build_oracle()
    build_devmath_host()
    build_em()
    build_compat_frame()


if __name__ == "__main__":
    build_all()
    print("built:", os.path.join(HERE, "libtgmm_em.so"))
