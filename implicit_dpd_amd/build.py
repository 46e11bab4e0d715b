"""Builds the engine library implicit_dpd_amd/libmdpd.so with hipcc for gfx950.

Every source in csrc/ is compiled by hipcc (HIP kernels and the host C++ of
the C ABI) in parallel, then linked into one shared library in-tree.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmdpd.so")
OBJ = os.path.join(HERE, "_build")
ARCH = os.environ.get("MDPD_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
SOURCES = ["k_step.hip", "k_pair.hip", "k_obs.hip", "k_mlp.hip", "k_config.hip", "engine.cpp", "params.cpp",
           "ext.cpp", "mlp.cpp", "config.cpp", "io.cpp", "rdf_noalloc.hip"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC] + FLAGS + lang + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    build_driver(verbose)
    return OUT


DRIVER = os.path.join(HERE, "bin", "mdpd_main")
CONFIG_MAKER = os.path.join(HERE, "bin", "config_maker")
# host programs over the C++ facade: the run driver (src/main.cpp) and the
# initial-configuration program (src/config_maker.cpp)
PROGRAMS = {DRIVER: os.path.join(HERE, "driver", "mdpd_main.cpp"),
            CONFIG_MAKER: os.path.join(HERE, "driver", "config_maker.cpp")}


def build_driver(verbose=True):
    """The stand-alone host programs (C++ facade over the C ABI), linked
    against the in-tree library with an rpath to it."""
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    for out, src in PROGRAMS.items():
        cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(os.path.dirname(HERE), "include"),
               src, "-o", out + ".tmp", "-L", HERE, "-lmdpd",
               "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host program build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(out + ".tmp", out)
        if verbose:
            print(f"built {out}")
    return DRIVER


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build()
    sys.exit(0)
