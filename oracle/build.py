"""Build the CPU oracle (test infrastructure) into oracle/lib/liboracle.so.

gcc, -O2, IEEE arithmetic: -ffp-contract=off and no fast-math, so the fp32
clock of the adapted grid rounds exactly as torch's fp32 tensors do.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "sdemc_oracle.c")
OUT = os.path.join(HERE, "lib", "liboracle.so")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
           "-fvisibility=hidden", "-D_GNU_SOURCE", "-I" + os.path.join(ROOT, "include"), SRC,
           "-o", tmp, "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
