"""Build libgvx.so (HIP kernels + C ABI, gfx950) and libgpu_voxels.so (C++ drop-in API).

    python gpu-voxels_amd/build.py            # both libraries, in-tree under gpu-voxels_amd/lib/

Every HIP source is compiled by hipcc for gfx950 only, with
-ffp-contract=off (voxel indices must round exactly as written) and packed
FP32 instructions disabled.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GVX_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + cmd[-1])
    return r.stderr


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    objs = []
    jobs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        for s in srcs:
            o = os.path.join(OBJ, os.path.basename(s) + ".o")
            objs.append(o)
            jobs.append(ex.submit(_run, [HIPCC] + HIP_FLAGS + ["-c", "-o", o, s]))
        for j in jobs:
            j.result()
    out = os.path.join(LIB, "libgvx.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out + ".tmp"] + objs)
    os.replace(out + ".tmp", out)
    return out


def build_cxx():
    srcs = sorted(glob.glob(os.path.join(HERE, "cpp", "*.cpp")))
    if not srcs:
        return None
    out = os.path.join(LIB, "libgpu_voxels.so")
    _run(["g++"] + CXX_FLAGS + ["-shared", "-o", out + ".tmp"] + srcs +
         ["-L", LIB, "-lgvx", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build_tests():
    """The C++ drop-in test program (restates the reference's Boost tests)."""
    src = os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp")
    out = os.path.join(LIB, "test_dropin")
    _run(["g++", "-O1", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", out + ".tmp", src,
          "-L", LIB, "-lgpu_voxels", "-lgvx", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build():
    a = build_hip()
    b = build_cxx()
    c = build_tests()
    return a, b, c


if __name__ == "__main__":
    print(build())
