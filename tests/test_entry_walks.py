"""The walk-skipping grid's entry trees on the CPU: a walk from a cell's
restricted tree (copies of the BVH's nodes below the deepest common node,
one-child chains skipped, children with nothing in reach never visited or
pushed; a second level for steps of up to three cells) finds the same hit,
at the same distance, as the walk from the BVH root, for every sampled ray.

tests/entry_walks.cpp builds the BVH, distance grid and entry grid with the
engine's own host code and restates the kernel's walk on the host; this test
compiles it with g++ and runs it on the benchmark heightfield (98,306
triangles), a small heightfield and a coarse head model (five nested closed
shells).  The GPU counterpart is test_entry_trees_change_nothing_on_the_full_
heightfield in test_gpu_parity.py (traces and tallies with and without the
grids)."""
import json
import shutil
import subprocess
from pathlib import Path

import pytest

from xmcml_amd import files, scenes

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "xmcml_amd" / "csrc"
SOURCES = [ROOT / "tests" / "entry_walks.cpp", CSRC / "xmc_bvh.cpp", CSRC / "xmc_entry.cpp", CSRC / "xmc_distance.cpp"]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    exe = tmp_path_factory.mktemp("entry") / "entry_walks"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", f"-I{CSRC}", f"-I{ROOT / 'include'}", "-o",
                    str(exe), *map(str, SOURCES), "-lpthread"], check=True)
    return exe


def _run(driver, spec, tmp_path, queries, seed):
    surf = tmp_path / "scene.surface"
    files.write_surface_file(spec, surf)
    out = subprocess.run([str(driver), str(surf), str(queries), str(seed)], check=True, capture_output=True,
                         text=True, timeout=600)
    return json.loads(out.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("name,make,queries", [
    ("heightfield", lambda: scenes.heightfield_scene(10), 150_000),
    ("heightfield-small", lambda: scenes.small_scene("heightfield", 10), 60_000),
    ("head-coarse", lambda: scenes.head_scene(10, level=3), 60_000),
])
def test_entry_walks_equal_root_walks(driver, tmp_path, name, make, queries):
    r = _run(driver, make(), tmp_path, queries, seed=11)
    print(name, r)
    assert r["mismatches"] == 0
    # every kind of query is exercised: both levels, cells with nothing in reach, hits
    assert r["level1"] > queries // 5 and r["level2"] > queries // 20
    assert r["no_entry"] > 0 and r["hits"] > queries // 50
    assert r["visits_entry"] < r["visits_root"]
