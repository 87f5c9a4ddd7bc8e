"""Build the CPU emulation of the library (test infrastructure): the
unmodified kernel and host sources compiled by g++ against the HIP subset in
tests/cpu_emu/include, into tests/cpu_emu/build/.  Never loaded by the
product (fastlem_amd._native loads only the gfx950 build)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "fastlem_amd", "csrc")
SOURCES = ["fl_kernels.hip", "fl_order.hip", "fl_create.hip", "fl_noalloc.hip", "fl_interp.hip", "fl_capi.cpp", "fl_setup.cpp", "fl_server.cpp"]


def lib_path(asan: bool = False) -> str:
    return os.path.join(HERE, "build", "libfastlem_emu_asan.so" if asan else "libfastlem_emu.so")


def build(asan: bool = False) -> str:
    out = lib_path(asan)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES] + [os.path.join(HERE, "emu.cpp")]
    newest = max(os.path.getmtime(p) for p in srcs + [os.path.join(HERE, "include", "hip", "hip_runtime.h")])
    if os.path.exists(out) and os.path.getmtime(out) >= newest:
        return out
    flags = ["-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", "-fno-fast-math",
             "-U_FORTIFY_SOURCE", "-D_FORTIFY_SOURCE=0", "-fno-omit-frame-pointer", "-w",
             "-I", os.path.join(HERE, "include"), "-I", os.path.join(ROOT, "include")]
    if asan:
        flags += ["-fsanitize=address", "-fsanitize=undefined"]
    cmd = ["g++"] + flags
    for s in srcs:
        if s.endswith(".hip"):
            cmd += ["-x", "c++", s, "-x", "none"]
        else:
            cmd.append(s)
    cmd += ["-o", out + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(asan="--asan" in sys.argv))
