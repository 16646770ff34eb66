"""Build the oracle's C restatement (test infrastructure only).

Compiled with -ffp-contract=off so the fixed-point draw arithmetic is the
literal sequence of IEEE operations written in flame_oracle.c.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"


def build() -> Path:
    LIB.parent.mkdir(exist_ok=True)
    srcs = [HERE / "flame_oracle.c", HERE / "jpeg_oracle.c"]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
           "-std=gnu11", *map(str, srcs), "-o", str(tmp), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stderr)
    tmp.replace(LIB)
    return LIB


if __name__ == "__main__":
    print(build())
