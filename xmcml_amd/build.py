"""Builds the engine library (hipcc, gfx950) and the CPU oracle (g++).

    python -m xmcml_amd.build             # library + oracle (+ reference build when present)
    python -m xmcml_amd.build --sanitize  # host-side ASan + UBSan builds for CPU test runs

The engine is one shared object, ``xmcml_amd/libxmcml_amd.so``, exporting
the C-ABI of include/xmcml.h.  Every HIP translation unit is compiled with
FMA contraction off (the float arithmetic is specified operation by
operation) and without packed FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "xmcml_amd"
CSRC = PKG / "csrc"
LIB = PKG / "libxmcml_amd.so"
REFERENCE = Path("/root/reference")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("XMC_OFFLOAD_ARCH", "gfx950")
SOURCES = ["xmc_transport.hip", "xmc_mlp.hip", "xmc_host.cpp", "xmc_bvh.cpp", "xmc_distance.cpp", "xmc_lattice.cpp",
           "xmc_entry.cpp", "xmc_io.cpp", "xmc_bias.hip"]
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], cwd=cwd, check=True)


def build_library() -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, f"-I{ROOT / 'include'}", *[CSRC / s for s in SOURCES], "-o", tmp]
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


# Host-code sanitizers (AddressSanitizer + UndefinedBehaviorSanitizer) for CPU test runs: each
# -fsanitize= goes to the host compilation only; the gfx950 code is built as usual.
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer",
            "-Xarch_host", "-shared-libsan", "-g"]
LIB_SAN = PKG / "libxmcml_amd_san.so"


def build_sanitized() -> None:
    """xmcml_amd/libxmcml_amd_san.so and oracle/liboracle_san.so with host-side
    ASan + UBSan (tests/run_sanitized.sh runs the CPU suite on them)."""
    tmp = LIB_SAN.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, *SAN_HOST, f"-I{ROOT / 'include'}", *[CSRC / s for s in SOURCES], "-o", tmp]
    _run(cmd)
    os.replace(tmp, LIB_SAN)
    _run(["make", "-s", "-C", ROOT / "oracle", "liboracle_san.so"])


def build_oracle() -> None:
    _run(["make", "-s", "-C", ROOT / "oracle", "liboracle.so"])
    if (REFERENCE / "base" / "xmcml" / "xmcml" / "mcml_kernel.cpp").exists():
        _run(["make", "-s", "-C", ROOT / "oracle", "ref", f"REF={REFERENCE}"])


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if "--sanitize" in argv:
        build_sanitized()
        return 0
    build_library()
    build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main())
