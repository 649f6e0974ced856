"""Build tests/abi/drive_tailor (C++20, linked against the in-tree
libtailor.so) and check that include/tailor.h compiles as C99.  Called by
the repository's build step, so the GPU tests find the binary built."""
from __future__ import annotations

import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
DRIVER = HERE / "drive_tailor"


def build(verbose: bool = False) -> Path:
    lib_dir = ROOT / "tailor_amd"
    subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", f"-I{ROOT / 'include'}",
                    "-o", str(DRIVER), str(HERE / "drive_tailor.cpp"), f"-L{lib_dir}", "-ltailor",
                    f"-Wl,-rpath,$ORIGIN/../../tailor_amd"], check=True)
    with tempfile.TemporaryDirectory() as tmp:
        c = Path(tmp) / "header.c"
        c.write_text('#include "tailor.h"\nint main(void) { tailor_ctx *c = 0; '
                     'return c == 0 ? 0 : tailor_eigenpair_count(c); }\n')
        subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", f"-I{ROOT / 'include'}",
                        "-c", str(c), "-o", str(Path(tmp) / "header.o")], check=True)
    if verbose:
        print("built", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
