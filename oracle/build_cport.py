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


if __name__ == "__main__":
    build("-q" not in sys.argv)
