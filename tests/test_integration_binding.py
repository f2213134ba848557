"""INTEGRATION.md's reference-side binding (LaunchAMD, the LaunchOMP
replacement) compiles as written against the reference's own
mcml_kernel_types.h, including its static_asserts that TimeInfo, DetectorInfo
and WeightIntegralTable have the layouts of xmc_time_info, xmc_detector_info
and xmc_weight_table (the binding shares those arrays by pointer).  Needs the
reference tree (build container only)."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF_KERNEL = Path("/root/reference/base/xmcml/xmcml")

pytestmark = pytest.mark.skipif(not (REF_KERNEL / "mcml_kernel_types.h").exists() or shutil.which("g++") is None,
                                reason="reference headers or g++ not present")


def binding_source() -> str:
    text = (ROOT / "INTEGRATION.md").read_text()
    blocks = re.findall(r"```cpp\n(.*?)```", text, re.S)
    src = next(b for b in blocks if "void LaunchAMD(" in b)
    return src


def test_binding_compiles_against_the_reference_headers(tmp_path):
    f = tmp_path / "launcher_gpu_amd.cpp"
    f.write_text(binding_source())
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-Wall", f"-I{REF_KERNEL}", f"-I{ROOT / 'include'}",
                        str(f)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_a_layout_mismatch_would_fail_to_compile(tmp_path):
    """The static_asserts bite: with one field of xmc_time_info reordered the
    block does not compile."""
    hdr = (ROOT / "include" / "xmcml.h").read_text()
    old = "    uint64_t numberOfPhotons;\n    double weight;\n    double targetWeight;\n} xmc_time_info;"
    assert old in hdr
    inc = tmp_path / "inc"
    inc.mkdir()
    (inc / "xmcml.h").write_text(hdr.replace(old, "    double weight;\n    uint64_t numberOfPhotons;\n"
                                                  "    double targetWeight;\n} xmc_time_info;"))
    f = tmp_path / "launcher_gpu_amd.cpp"
    f.write_text(binding_source())
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", f"-I{REF_KERNEL}", f"-I{inc}", str(f)],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "TimeInfo.numberOfPhotons" in r.stderr
