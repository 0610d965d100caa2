"""Build libgd2.so (HIP, gfx950) with hipcc, and with g++ the LibTorch adjoint
module gd2_torch and the production-run driver gd2_production (both host C++
over the libgd2 C ABI), all in-tree."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
SO = os.path.join(HERE, "libgd2.so")
SOURCES = ["gd2_api.hip", "gd2_variants.hip", "gd2_mhd_vjp.hip", "gd2_io.cpp"]
# per-source code-generation options: the MHD stage VJP kernels take the
# machine scheduler's ILP strategy (measured faster there, slower on the
# Euler kernels)
SOURCE_FLAGS = {"gd2_mhd_vjp.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    # contract a*b+c only within one source expression: the rounding of every
    # expression is then fixed by the source, not by later code motion
    "-ffp-contract=on",
    "--offload-arch=gfx950",
    # no packed-FP32 instructions (device code is FP64 throughout)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function",
]


TORCH_SRC = "gd2_torch.cpp"


def torch_module_path() -> str:
    import sysconfig
    return os.path.join(HERE, "gd2_torch" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_torch_module(verbose: bool = False) -> str:
    """g++ build of the pybind11/LibTorch module, linked against libgd2.so."""
    import sysconfig

    import torch
    from torch.utils import cpp_extension as ce

    out = torch_module_path()
    tlib = ce.library_paths()[0]
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-w",
           "-DTORCH_EXTENSION_NAME=gd2_torch", "-DTORCH_API_INCLUDE_EXTENSION_H",
           f"-D_GLIBCXX_USE_CXX11_ABI={int(torch._C._GLIBCXX_USE_CXX11_ABI)}",
           *[f"-I{p}" for p in ce.include_paths()], f"-I{sysconfig.get_paths()['include']}",
           "-D__HIP_PLATFORM_AMD__=1", "-I", "/opt/rocm/include",
           "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, TORCH_SRC),
           f"-L{tlib}", "-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python", f"-Wl,-rpath,{tlib}",
           f"-L{HERE}", "-lgd2", "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64", "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


PRODUCTION_SRC = "gd2_production.cpp"
PRODUCTION = os.path.join(HERE, "gd2_production")


def build_production(verbose: bool = False) -> str:
    """g++ build of the production-run driver, an executable over libgd2.so."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__=1", "-I", "/opt/rocm/include",
           "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, PRODUCTION_SRC),
           f"-L{HERE}", "-lgd2", "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,/opt/rocm/lib", "-o", PRODUCTION + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(PRODUCTION + ".tmp", PRODUCTION)
    return PRODUCTION


def build(verbose: bool = False, extra=()) -> str:
    """Compile each source to an object in parallel (in a build folder next to
    the library), then link libgd2.so."""
    from concurrent.futures import ThreadPoolExecutor

    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    compile_flags = [f for f in FLAGS if f != "-shared"]
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *compile_flags, *SOURCE_FLAGS.get(src, []), *extra, "-I", os.path.join(ROOT, "include"),
                     "-c", os.path.join(CSRC, src), "-o", obj])

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for f in [ex.submit(run, c) for c in cmds]:
            f.result()
    link = [HIPCC, *FLAGS, *extra, *objs, "-o", SO + ".tmp"]
    run(link)
    os.replace(SO + ".tmp", SO)
    return SO


if __name__ == "__main__":
    build(verbose=True, extra=[a for a in sys.argv[1:] if a != "--torch"])
    build_production(verbose=True)
    if "--torch" in sys.argv[1:]:
        build_torch_module(verbose=True)
