"""Build recipe for libcdixx_amd.so (HIP kernels for gfx950 + C ABI + host C++).

Run as ``python cdixx_amd/build.py`` or through ``__graft_entry__.build()`` (not
through the package, whose import needs the built library).
Objects go to cdixx_amd/_build/, the library to cdixx_amd/libcdixx_amd.so.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcdixx_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU ops.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-Wall",
             "-Wno-unused-function", "-Wno-unused-variable",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
              f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]

# (source, extra defines, object name); passes.hip builds once per precision
# and pass direction so its template instantiations compile in parallel
UNITS = [
    ("passes.hip", ["-DCDX_PASS_T=float", "-DCDX_PASS_ROW=0", "-DCDX_PASS_HOST"], "passes_col_f32.o"),
    ("passes.hip", ["-DCDX_PASS_T=double", "-DCDX_PASS_ROW=0"], "passes_col_f64.o"),
    ("passes.hip", ["-DCDX_PASS_T=float", "-DCDX_PASS_ROW=1"], "passes_row_f32.o"),
    ("passes.hip", ["-DCDX_PASS_T=double", "-DCDX_PASS_ROW=1"], "passes_row_f64.o"),
    ("kernels.hip", [], "kernels.o"),
    ("abi.hip", [], "abi.o"),
    ("abi_extra.hip", [], "abi_extra.o"),
    ("batch.hip", [], "batch.o"),
    ("unet.hip", [], "unet.o"),
    ("preprocess.hip", [], "preprocess.o"),
    ("ptycho.hip", [], "ptycho.o"),
    ("config.cpp", [], "config.o"),
    ("cdi.cpp", [], "cdi.o"),
    ("imgio.cpp", [], "imgio.o"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _compile(unit):
    src, defs, objname = unit
    path = os.path.join(CSRC, src)
    obj = os.path.join(OUT, objname)
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + defs + ["-c", path, "-o", obj]
    else:
        cmd = ["g++"] + HOST_FLAGS + defs + ["-c", path, "-o", obj]
    _run(cmd)
    return obj


def build(verbose=True, jobs=None):
    os.makedirs(OUT, exist_ok=True)
    units = [u for u in UNITS if os.path.exists(os.path.join(CSRC, u[0]))]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, units))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,-z,defs", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # the command-line driver (cdi_run equivalent)
    cli_src = os.path.join(CSRC, "cdi_run.cpp")
    if os.path.exists(cli_src):
        _run(["g++"] + HOST_FLAGS + ["-o", os.path.join(HERE, "cdi_run"), cli_src,
                                    f"-L{HERE}", "-lcdixx_amd", f"-Wl,-rpath,{HERE}",
                                    "-Wl,-rpath,$ORIGIN"])
    # a test program: a caller of include/cdi.hpp shaped like the
    # reference's drivers (run by tests/test_gpu_dropin.py)
    caller = os.path.join(ROOT, "tests", "dropin", "cdi_caller.cpp")
    if os.path.exists(caller):
        _run(["g++"] + HOST_FLAGS + [f"-I{os.path.join(ROOT, 'include')}", "-o", caller[:-4], caller,
                                    f"-L{HERE}", "-lcdixx_amd", f"-Wl,-rpath,{HERE}",
                                    "-Wl,-rpath,$ORIGIN/../../cdixx_amd"])
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
