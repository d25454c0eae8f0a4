"""Builds the FP64 C port (oracle/cport/libbcport.so) with gcc -O3 -fopenmp."""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cport", "stagger_port.c")
OUT = os.path.join(HERE, "cport", "libbcport.so")


def build(verbose=True):
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", SRC, "-o", OUT + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


ASAN_OUT = os.path.join(HERE, "cport", "asan", "libbcport_asan.so")


def build_asan(verbose=False):
    """The port with AddressSanitizer and UndefinedBehaviorSanitizer (test
    build, loaded under gcc's ASan runtime by the CPU sanitizer test)."""
    os.makedirs(os.path.dirname(ASAN_OUT), exist_ok=True)
    cmd = ["gcc", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fopenmp", "-fPIC", "-shared", SRC, "-o", ASAN_OUT + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(ASAN_OUT + ".tmp", ASAN_OUT)
    return ASAN_OUT


if __name__ == "__main__":
    if "--asan" in sys.argv:
        build_asan("-q" not in sys.argv)
    else:
        build("-q" not in sys.argv)
