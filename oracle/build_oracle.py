"""Build the oracle (CPU FP64 restatement of the reference residual and of
the vortex-lattice / beam coupling path).

Test infrastructure only: produces oracle/liboracle_rans.so with g++ and
-ffp-contract=off so that every operation rounds as written.
Run: python oracle/build_oracle.py
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_rans.so")


def build():
    srcs = [os.path.join(HERE, "oracle_rans.cpp"), os.path.join(HERE, "oracle_vlm.cpp")]
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math"] + srcs + [
        "-o", tmp]
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
