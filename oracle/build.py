"""Build the oracle's C restatement (test infrastructure only).

Compiled with -ffp-contract=off so the fixed-point draw arithmetic is the
literal sequence of IEEE operations written in flame_oracle.c.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"


FAST = HERE / "_build" / "liboracle_f32.so"


def _gcc(cmd: list[str], out: Path) -> None:
    tmp = out.with_suffix(".so.tmp")
    r = subprocess.run([*cmd, "-o", str(tmp), "-lm"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stderr)
    tmp.replace(out)


def build() -> Path:
    LIB.parent.mkdir(exist_ok=True)
    srcs = [HERE / "flame_oracle.c", HERE / "jpeg_oracle.c"]
    _gcc(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-std=gnu11",
          *map(str, srcs)], LIB)
    build_f32(FAST)
    return LIB


def build_f32(out: Path) -> Path:
    """the host-core baseline: the same render loop in single precision,
    built for the machine that runs it (-march=native)"""
    out.parent.mkdir(parents=True, exist_ok=True)
    _gcc(["gcc", "-O3", "-march=native", "-fPIC", "-shared", "-fopenmp", "-DORC_REAL=float",
          "-fsingle-precision-constant", "-std=gnu11", str(HERE / "flame_oracle.c")], out)
    return out


if __name__ == "__main__":
    print(build())
