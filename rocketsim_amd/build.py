"""Build recipes for the native library (librsg.so, HIP for gfx950).

The library is built in-tree (rocketsim_amd/lib/) so it travels with the
source tree; nothing is installed into site-packages.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB_PATH = os.path.join(LIBDIR, "librsg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off keeps a*b+c as two rounded operations like the reference's
# x86-64 build; packed FP32 instructions are disabled (some of their forms have
# given wrong results on gfx950 without an error).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# Translation units of librsg.so: the C ABI (host code) and one file per
# kernel family, compiled in parallel and linked into one shared library.
NATIVE_SOURCES = ["rsg_batch.hip", "rsg_k_plain.hip", "rsg_k_lds.hip", "rsg_k_coop.hip", "rsg_k_pred.hip",
                  "rsg_k_diff.hip", "rsg_k_events.hip", "rsg_k_status.hip"]


def build_native(verbose=True, extra_flags=(), out_path=None, obj_dir=None):
    """Compile every translation unit of NATIVE_SOURCES for gfx950 (in
    parallel) and link librsg.so.  `extra_flags` (e.g. -D switches) and
    `out_path` / `obj_dir` let a caller build a variant beside the product."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIBDIR, exist_ok=True)
    out_path = out_path or LIB_PATH
    obj_dir = obj_dir or os.path.join(LIBDIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    inc = "-I" + os.path.join(ROOT, "include")

    def compile_one(name):
        src = os.path.join(CSRC, name)
        obj = os.path.join(obj_dir, name.replace(".hip", ".o"))
        cmd = [HIPCC, *HIP_FLAGS, *extra_flags, inc, "-c", src, "-o", obj]
        if verbose:
            print("[build]", " ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        return name, obj, r

    jobs = min(len(NATIVE_SOURCES), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(compile_one, NATIVE_SOURCES))
    for name, _, r in results:
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {name}")
    tmp = out_path + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", *[o for _, o, _ in results], "-o", tmp]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking librsg.so")
    os.replace(tmp, out_path)
    return out_path


TORCH_EXT_PATH = os.path.join(LIBDIR, "rsg_torch.so")


def build_torch_ext(verbose=True):
    """The LibTorch autograd Function module (csrc/rsg_torch.cpp), linked to
    librsg.so beside it.  Host C++ only: it calls the C ABI, which launches the
    HIP kernels."""
    import sysconfig

    import torch
    from torch.utils import cpp_extension
    os.makedirs(LIBDIR, exist_ok=True)
    src = os.path.join(CSRC, "rsg_torch.cpp")
    tmp = TORCH_EXT_PATH + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", src, "-o", tmp,
           "-I" + os.path.join(ROOT, "include"), "-I" + sysconfig.get_paths()["include"],
           "-D_GLIBCXX_USE_CXX11_ABI=" + str(int(torch._C._GLIBCXX_USE_CXX11_ABI)),
           "-DTORCH_EXTENSION_NAME=rsg_torch", "-DTORCH_API_INCLUDE_EXTENSION_H"]
    cmd += ["-I" + p for p in cpp_extension.include_paths()]
    cmd += ["-L" + p for p in cpp_extension.library_paths()]
    cmd += ["-L" + LIBDIR, "-lrsg", "-Wl,-rpath,$ORIGIN", "-ltorch", "-ltorch_cpu", "-lc10", "-ltorch_python"]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building rsg_torch.so")
    os.replace(tmp, TORCH_EXT_PATH)
    return TORCH_EXT_PATH


if __name__ == "__main__":
    build_native()
    build_torch_ext()
