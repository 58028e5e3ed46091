"""Build the native libraries of libfusion_amd in-tree.

    liblf_synth.so    host C++ (g++ -fopenmp): seeded synthetic RGB-D scenes
    libfusion_amd.so  HIP for gfx950 (hipcc): kernels, C-ABI and the C++ host classes

Every HIP translation unit is compiled with
  -ffp-contract=off     each float operation rounded on its own (the rule the
                        oracle is built with, so both sides round identically)
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
The build always recompiles (it is the "does it build" check).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")

HIP_SOURCES = [
    "common.hip",
    "map.hip",
    "raycast.hip",
    "mesh.hip",
    "imgproc.hip",
    "reduce.hip",
    "render.hip",
    "system.cpp",
    "map_io.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LF_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), r.stdout))
    return r.stdout


def build_synth():
    out = os.path.join(HERE, "liblf_synth.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-o", out,
          os.path.join(CSRC, "synth.cpp")])
    return out


def build_hip(jobs=8):
    os.makedirs(OBJ, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC] + HIP_FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        _run(cmd)
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    out = os.path.join(HERE, "libfusion_amd.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


def build_oracle():
    _run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liblf_oracle.so")


def build_all(hip=True):
    outs = [build_synth(), build_oracle()]
    if hip:
        outs.append(build_hip())
    return outs


if __name__ == "__main__":
    hip = "--no-hip" not in sys.argv
    for o in build_all(hip=hip):
        print(o)
