"""Build recipe for libeqbf_hip.so (HIP for gfx950) and the eqtlbma_bf and
eqtlbma_hm host programs.  Objects go to eqtlbma_amd/_build/; the shared library and the
program are written next to this file so they travel with the source tree.

    python -m eqtlbma_amd.build            # library + programs
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "libeqbf_hip.so")
CLI = os.path.join(HERE, "eqtlbma_bf")
HOST_HM = os.path.join(HERE, "host_hm")
CLI_HM = os.path.join(HERE, "eqtlbma_hm")
HOST_AVG = os.path.join(HERE, "host_avg")
CLI_AVG = os.path.join(HERE, "eqtlbma_avg_bfs")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("EQBF_ARCH", "gfx950")
# packed FP32 instructions are disabled for every HIP object
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-I" + INCLUDE]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; drop that note
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l
                    and l.strip() != "1 warning generated." and "warning generated" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_lib(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip"))) + sorted(glob.glob(os.path.join(CSRC, "*.cpp")))
    objs = []

    def comp(src):
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        if src.endswith(".cpp"):  # host-only C++
            _run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + INCLUDE, "-c", src, "-o", obj])
        else:
            _run([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(comp, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_cli():
    srcs = sorted(glob.glob(os.path.join(HOST, "*.cpp")))
    if not srcs:
        return None
    tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, "-o", tmp] + srcs +
         ["-L" + HERE, "-leqbf_hip", "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib",
          "-lz", "-lpthread"])
    os.replace(tmp, CLI)
    return CLI


def build_hm():
    srcs = sorted(glob.glob(os.path.join(HOST_HM, "*.cpp"))) + [os.path.join(HOST, "util.cpp")]
    tmp = CLI_HM + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, "-o", tmp] + srcs +
         ["-L" + HERE, "-leqbf_hip", "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib",
          "-lz"])
    os.replace(tmp, CLI_HM)
    return CLI_HM


def build_avg():
    srcs = sorted(glob.glob(os.path.join(HOST_AVG, "*.cpp"))) + [os.path.join(HOST, "util.cpp")]
    tmp = CLI_AVG + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, "-o", tmp] + srcs +
         ["-L" + HERE, "-leqbf_hip", "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib",
          "-lz"])
    os.replace(tmp, CLI_AVG)
    return CLI_AVG


def build(jobs=8):
    build_lib(jobs)
    build_cli()
    build_hm()
    build_avg()


if __name__ == "__main__":
    build(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    print("built", LIB)
