"""Build recipe for the C oracle (test infrastructure): oracle/liboracle_ros2.so.

Portable x86-64-v3 code (AVX2 + FMA) so the same object runs on any host of
the GPU boxes; OpenMP for the one-column-per-thread baseline.
"""
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "ros2_ref.c"
OUT = HERE / "liboracle_ros2.so"


def build():
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c11",
           str(SRC), "-o", str(OUT), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stderr}")
    print(f"built {OUT}", file=sys.stderr)
    return OUT


if __name__ == "__main__":
    build()
