"""Compile the CPU oracle (test infrastructure) with gcc into oracle/_build/."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "liboracle.so")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = ["gcc", "-O2", "-std=c99", "-fopenmp", "-Wall", "-Wextra", "-fPIC", "-shared",
           os.path.join(HERE, "wk_oracle.c"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
