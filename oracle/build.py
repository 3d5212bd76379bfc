"""Build recipe for the oracle (test infrastructure): oracle/liboracle.so.

g++ only; OpenMP over x-planes stands in for the reference's rayon loop (cpu.rs:60, 120).
-ffp-contract=off keeps the reference's operation order (no fused multiply-adds).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = HERE
LIB = os.path.join(OUT_DIR, "liboracle.so")


def build():
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-Wall", "-Wno-missing-field-initializers",
           os.path.join(HERE, "fdtd_oracle.cpp"), "-o", tmp]
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
