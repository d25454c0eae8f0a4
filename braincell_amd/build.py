"""Builds the HIP library ``braincell_amd/_lib/libbraincell_amd.so`` for gfx950.

``python -m braincell_amd.build`` (or ``__graft_entry__.build()``) compiles
every ``csrc/*.hip`` source with hipcc into one shared library in-tree.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIBNAME = "libbraincell_amd.so"
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
ARCH = os.environ.get("BC_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    # keep packed FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


# Per-source flags: the closure kernels take the VGPR form of MFMA (the AGPR
# form costs one v_accvgpr_read per accumulator element read back).
EXTRA_FLAGS = {"bc_closure.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def asan_library_path() -> str:
    return os.path.join(LIBDIR, "asan", "libbraincell_amd_asan.so")


def build_asan(verbose: bool = False) -> str:
    """Test build of the same sources with AddressSanitizer on the host code
    (descriptor validation, launch set-up, graph handling): -O0 -g, device
    code compiled normally.  Loaded under the clang ASan runtime by the CPU
    sanitizer test; never used for device work."""
    from concurrent.futures import ThreadPoolExecutor

    out_dir = os.path.dirname(asan_library_path())
    os.makedirs(out_dir, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    flags = ["-O0", "-g", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
    cmds, objs = [], []
    for src in sources():
        obj = os.path.join(out_dir, os.path.basename(src).replace(".hip", ".o"))
        cmds.append([hipcc, *flags, "-c", "-I", CSRC, "-I", INCLUDE, src, "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as ex:
        for r in list(ex.map(lambda c: subprocess.run(c, check=True, capture_output=not verbose), cmds)):
            pass
    out = asan_library_path()
    subprocess.run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Xarch_host", "-fsanitize=address", *objs,
                    "-o", out + ".tmp"], check=True, capture_output=not verbose)
    os.replace(out + ".tmp", out)
    return out


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def library_path() -> str:
    return os.path.join(LIBDIR, LIBNAME)


def build(verbose: bool = True) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    from concurrent.futures import ThreadPoolExecutor

    cmds, objs = [], []
    for src in sources():
        obj = os.path.join(LIBDIR, os.path.basename(src).replace(".hip", ".o"))
        extra = EXTRA_FLAGS.get(os.path.basename(src), [])
        cmds.append([hipcc, *[f for f in HIPCC_FLAGS if f != "-shared"], *extra, "-c", "-I", CSRC, "-I", INCLUDE, src,
                     "-o", obj])
        objs.append(obj)
        if verbose:
            print(" ".join(cmds[-1]), flush=True)
    # one hipcc per source, concurrently (each is single-threaded)
    with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as ex:
        for r in list(ex.map(lambda c: subprocess.run(c, check=True), cmds)):
            pass
    out = library_path()
    tmp = out + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    if "--asan" in sys.argv:
        print(build_asan(verbose="-q" not in sys.argv))
    else:
        build(verbose="-q" not in sys.argv)
