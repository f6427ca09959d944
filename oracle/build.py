"""Build recipe for the CPU oracle (test infrastructure only).

Compiles oracle/ngi_oracle.cpp into oracle/_build/liboracle.so with g++.
The reference renderer itself cannot be compiled here (it needs Embree 2,
TBB, yaml-cpp, Assimp, glm, Boost, FreeImage, ctemplate and Eigen3), so no
oracle/_ref build exists.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "liboracle.so")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra",
           os.path.join(HERE, "ngi_oracle.cpp"), "-o", tmp, "-lpthread"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
    sys.exit(0)
