"""Builds the oracle (test infrastructure) into oracle/liboracle.so with g++.

-ffp-contract=off and no fast-math: every f32 product and sum is rounded
separately, as in the reference (Rust does not contract floating point).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
SRC = [os.path.join(HERE, "voidray_oracle.cpp"), os.path.join(HERE, "nrc_oracle.cpp")]


def build():
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off",
           "-fno-fast-math", "-mno-fma", *SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
