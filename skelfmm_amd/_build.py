"""Compile ``libskelfmm_hip.so`` for gfx950 in-tree (``python -m skelfmm_amd._build``)."""

from __future__ import annotations

import hashlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = [os.path.join(HERE, "csrc", f) for f in ("skelfmm_hip.hip", "skelfmm_build.hip", "skelfmm_tree.hip")]
OUT = os.path.join(HERE, "libskelfmm_hip.so")
# the PyTorch operator library (host C++ over the C ABI, linked against OUT)
TORCH_SRC = os.path.join(HERE, "csrc", "torch_ops.cpp")
TORCH_OUT = os.path.join(HERE, "libskelfmm_torch.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-shared",
    "-std=c++17",
    "-munsafe-fp-atomics",
    # packed FP32 VALU ops stay off (the compiler would otherwise form them from float code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = os.path.join(rocm, "bin", "hipcc")
    return cand if os.path.exists(cand) else "hipcc"


# marker in the built file that precedes the source hash (found without loading the library)
BUILD_TAG = b"skelfmm-build-id:"


def source_files():
    """Every file the library is compiled from: the HIP sources, their headers, the C ABI header."""
    files = [os.path.join(HERE, "csrc", f) for f in os.listdir(os.path.join(HERE, "csrc"))
             if f.endswith((".hip", ".h", ".cpp"))]
    files += [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")]
    return sorted(files)


def source_hash():
    """16 hex digits of SHA-256 over the library's sources and compile flags."""
    h = hashlib.sha256()
    for path in source_files():
        h.update(os.path.relpath(path, ROOT).encode() + b"\0")
        with open(path, "rb") as fh:
            h.update(fh.read())
    h.update(" ".join(HIPCC_FLAGS).encode())
    return h.hexdigest()[:16]


def built_hash(path=OUT):
    """The source hash embedded in a built library file, or None (no file, no tag)."""
    if not os.path.exists(path):
        return None
    with open(path, "rb") as fh:
        data = fh.read()
    at = data.find(BUILD_TAG)
    if at < 0:
        return None
    return data[at + len(BUILD_TAG):at + len(BUILD_TAG) + 16].decode("ascii", errors="replace")


def build_library(verbose=False):
    """Compile every translation unit to an object in parallel, then link the shared library."""
    import shutil
    import tempfile
    from concurrent.futures import ThreadPoolExecutor

    tag = f'-DSKF_SOURCE_HASH="{source_hash()}"'
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    objdir = tempfile.mkdtemp(prefix="skelfmm_objs_")
    try:
        objs = [os.path.join(objdir, os.path.basename(src) + ".o") for src in SRC]

        def compile_one(pair):
            src, obj = pair
            cmd = [hipcc()] + compile_flags + [tag, "-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            return subprocess.run(cmd, capture_output=True, text=True)

        with ThreadPoolExecutor(max_workers=len(SRC)) as pool:
            results = list(pool.map(compile_one, zip(SRC, objs)))
        for src, res in zip(SRC, results):
            if res.returncode != 0:
                sys.stderr.write(res.stdout + res.stderr)
                raise RuntimeError(f"hipcc failed to compile {os.path.basename(src)}")
        tmp = OUT + ".tmp"
        cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,-soname,libskelfmm_hip.so"] + objs + \
            ["-lhiprtc", "-o", tmp]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("linking libskelfmm_hip.so failed")
        os.replace(tmp, OUT)
    finally:
        shutil.rmtree(objdir, ignore_errors=True)
    return OUT


def build_torch_ops(verbose=False):
    """Compile the PyTorch operator library (``torch.ops.skelfmm.apply``) with the host compiler
    against PyTorch's headers and libraries; it links the HIP library beside it (rpath $ORIGIN)
    and carries the same source hash."""
    import torch
    from torch.utils import cpp_extension

    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    inc = cpp_extension.include_paths()
    libdir = cpp_extension.library_paths()[0]
    tmp = TORCH_OUT + ".tmp"
    cmd = (["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-D_GLIBCXX_USE_CXX11_ABI={int(torch._C._GLIBCXX_USE_CXX11_ABI)}",
            "-DUSE_ROCM", "-D__HIP_PLATFORM_AMD__", f'-DSKF_SOURCE_HASH="{source_hash()}"']
           + [f"-I{p}" for p in inc] + [f"-I{rocm}/include", "-I", os.path.join(ROOT, "include"), TORCH_SRC,
                                        f"-L{libdir}", "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip",
                                        f"-L{HERE}", "-lskelfmm_hip", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{libdir}",
                                        "-o", tmp])
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("compiling libskelfmm_torch.so failed")
    os.replace(tmp, TORCH_OUT)
    return TORCH_OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_torch_ops(verbose=True))
