"""The transport kernel's BVH walk, its own source compiled for the host
(tests/walk_host.cpp), against brute force over every triangle: random rays,
half of them unlimited (the glass-layer walk), on the 98,306-triangle
heightfield and the small heightfield.  Built three ways -- with ASan + UBSan,
and with locals initialised to a pattern or to zero (a read of an
uninitialised value would change the result between the two) -- every build
must report no sanitizer error and no disagreement, for walk() and for the
walk_begin / walk_visit loop."""
import ctypes as C
import shutil
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as ol
from xmcml_amd import scenes
from xmcml_amd.abi import InputHandle

ROOT = Path(__file__).resolve().parents[1]
CLANG = "/opt/rocm/llvm/bin/clang++"
pytestmark = pytest.mark.skipif(not Path(CLANG).exists(), reason="clang++ of the ROCm toolchain not present")

DRIVER = r'''
import ctypes as C, sys, numpy as np
sys.path[:0] = [sys.argv[2], sys.argv[2] + "/tests"]
from xmcml_amd import scenes
from xmcml_amd.abi import InputHandle
lib = C.CDLL(sys.argv[1])
lib.walk_compare.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
n = int(sys.argv[3])
for spec, lo, hi in ((scenes.heightfield_scene(10), [-19, -19, -1.5], [19, 19, 3.0]),
                     (scenes.small_scene("heightfield", 10), [-4, -4, -1.0], [4, 4, 3.0])):
    inp = InputHandle(spec)
    rng = np.random.default_rng(7)
    o = rng.uniform(lo, hi, size=(n, 3)); d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    s = rng.exponential(0.3, size=n); s[rng.random(n) < 0.5] = 0.0
    st = np.zeros(3, np.int32)
    assert lib.walk_compare(C.cast(inp.c.surface, C.c_void_p), inp.c.numberOfSurfaces, o.ctypes.data, d.ctypes.data,
                            s.ctypes.data, n, st.ctypes.data) == 0
    print(int(st[0]), int(st[1]), int(st[2]), flush=True)
'''


def walk_region() -> str:
    lines = (ROOT / "xmcml_amd" / "csrc" / "xmc_transport.hip").read_text().split("\n")
    start = next(i for i, l in enumerate(lines) if l.startswith("constexpr float kInf"))
    end = next(i for i, l in enumerate(lines) if l.startswith("// photon physics"))
    return "\n".join(lines[start:end - 1]) + "\n"


@pytest.mark.parametrize("mode", ["sanitizers", "pattern", "zero"])
def test_kernel_walk_on_host_equals_brute_force(tmp_path, mode):
    (tmp_path / "walk_region.inc").write_text(walk_region())
    shutil.copy(ROOT / "tests" / "walk_host.cpp", tmp_path / "walk_host.cpp")
    flags = ["-O2", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-shared-libsan"] \
        if mode == "sanitizers" else ["-O3", f"-ftrivial-auto-var-init={mode}"]
    so = tmp_path / "walk_host.so"
    subprocess.run([CLANG, *flags, "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                    f"-I{ROOT / 'xmcml_amd' / 'csrc'}", f"-I{ROOT / 'include'}", str(tmp_path / "walk_host.cpp"),
                    str(ROOT / "xmcml_amd" / "csrc" / "xmc_bvh.cpp"), str(ROOT / "xmcml_amd" / "csrc" / "xmc_lattice.cpp"),
                    "-o", str(so)], check=True)
    env = {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1", "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
           "PATH": "/usr/bin:/bin"}
    if mode == "sanitizers":
        env["LD_PRELOAD"] = next(Path("/opt/rocm/llvm/lib/clang").glob("*/lib/linux/libclang_rt.asan-x86_64.so")).as_posix()
    r = subprocess.run([sys.executable, "-c", DRIVER, str(so), str(ROOT), "4000"], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr
    rows = [list(map(int, l.split())) for l in r.stdout.strip().split("\n")]
    for found, bad_walk, bad_visit in rows:
        assert found > 1000 and bad_walk == 0 and bad_visit == 0, rows


def lattice_scenes():
    """Heightfield scenes under the surface rule: the 98,306-triangle
    heightfield (three lattice surfaces and a rectangle), the small one, and
    the small one with a hole (cells with one or no triangle)."""
    holey = scenes.small_scene("heightfield", 10)
    v, t = holey.surfaces[1]
    keep = np.ones(len(t), bool)
    keep[100:160] = False  # thirty cells without triangles
    keep[301:340:2] = False  # and twenty with one
    holey.surfaces[1] = (v, t[keep])
    return {"heightfield": (scenes.heightfield_scene(10), 20.5),
            "small": (scenes.small_scene("heightfield", 10), 5.5),
            "holey": (holey, 5.5)}


@pytest.fixture(scope="module")
def rule_lib(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("rule")
    (tmp / "walk_region.inc").write_text(walk_region())
    shutil.copy(ROOT / "tests" / "walk_host.cpp", tmp / "walk_host.cpp")
    so = tmp / "rule_host.so"
    subprocess.run([CLANG, "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                    f"-I{ROOT / 'xmcml_amd' / 'csrc'}", f"-I{ROOT / 'include'}", str(tmp / "walk_host.cpp"),
                    str(ROOT / "xmcml_amd" / "csrc" / "xmc_bvh.cpp"), str(ROOT / "xmcml_amd" / "csrc" / "xmc_lattice.cpp"),
                    "-o", str(so)], check=True)
    lib = C.CDLL(str(so))
    lib.rule_query.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 3 + [C.c_int] + [C.c_void_p] * 4
    return lib


def oracle_rule(spec, o, d, s):
    inp = InputHandle(spec)
    n = len(s)
    found = np.zeros(n, np.int32)
    dist = np.zeros(n)
    surf = np.zeros(n, np.int32)
    tri = np.zeros(n, np.int32)
    nrm = np.zeros(3 * n)
    rc = ol.oracle().xmco_query(inp.ptr, ol.dptr(o), ol.dptr(d), ol.dptr(s), n, 32, 2, ol.iptr(found),
                                ol.dptr(dist), ol.iptr(surf), ol.iptr(tri), ol.dptr(nrm))
    return rc, found, dist, surf, tri


@pytest.mark.parametrize("kind", ["heightfield", "small", "holey"])
def test_kernel_surface_rule_on_host_equals_oracle(rule_lib, kind):
    """The kernel's rule_hit (its own source, built for the host, with the
    engine's lattice builder) against the 32-bit oracle's surface rule on
    random rays through the layers: bit-identical hits, distances and
    triangles, for steps inside one cell, across up to 4 x 4 cells, longer
    (the walk), and unlimited."""
    spec, half = lattice_scenes()[kind]
    rng = np.random.default_rng(17)
    n = 40000
    o = np.stack([rng.uniform(-half, half, n), rng.uniform(-half, half, n), rng.uniform(-1.6, 3.6, n)], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = o.astype(np.float32).astype(np.float64)
    d = d.astype(np.float32).astype(np.float64)
    s = rng.exponential(0.2, n) * np.where(rng.random(n) < 0.1, 10.0, 1.0)
    s[::7] = 0.0  # unlimited
    inp = InputHandle(spec)
    found = np.zeros(n, np.int32)
    dist = np.zeros(n)
    surf = np.zeros(n, np.int32)
    idx = np.zeros(n, np.int32)
    assert rule_lib.rule_query(C.cast(inp.c.surface, C.c_void_p), inp.c.numberOfSurfaces, o.ctypes.data,
                               d.ctypes.data, s.ctypes.data, n, found.ctypes.data, dist.ctypes.data,
                               surf.ctypes.data, idx.ctypes.data) == 0
    rc, f2, d2, s2, t2 = oracle_rule(spec, o, d, s)
    assert rc == 0
    assert found.sum() > 3000
    assert np.array_equal(found, f2)
    assert np.array_equal(dist, d2)
    assert np.array_equal(surf, s2)
    assert np.array_equal(idx, t2)
