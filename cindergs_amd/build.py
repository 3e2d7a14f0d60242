"""In-tree build of the gfx950 HIP library (and, for tests, the CPU oracle)."""
from __future__ import annotations

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _make(directory: str, always: bool) -> None:
    args = ["make", "-C", directory, "-j8"]
    if always:
        args.insert(1, "-B")
    subprocess.run(args, check=True)


def build_library(always: bool = True) -> str:
    """Compile every HIP source for gfx950 into cindergs_amd/libcindergs_hip.so."""
    _make(os.path.join(ROOT, "cindergs_amd", "csrc"), always)
    return os.path.join(ROOT, "cindergs_amd", "libcindergs_hip.so")


def build_oracle(always: bool = True) -> str:
    """Compile the CPU oracle (test infrastructure) into oracle/liboracle.so."""
    _make(os.path.join(ROOT, "oracle"), always)
    return os.path.join(ROOT, "oracle", "liboracle.so")
