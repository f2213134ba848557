"""Writes tests/golden/tape_traces.npz: per-step traces of photons 0..7 of
seed 99 on the four small scenes (slab, glass, mismatch, heightfield), walked
by the reference's own sources (oracle/_ref/libxmcml_ref_tape.so, built by
oracle/Makefile) while they draw the GPU engine's 24-bit uniforms -- the top
24 bits of the same per-photon MCG59 states.  The GPU test compares the
engine's traces against these without the reference tree.  Run in a
container that has the reference tree:

    make -C oracle ref && python tests/golden/make_tape_traces.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT), str(ROOT / "tests")]

import oracle_lib as ol  # noqa: E402
from xmcml_amd import scenes  # noqa: E402
from xmcml_amd.abi import InputHandle  # noqa: E402

SEED, PHOTONS = 99, 8
SCENES = {
    "slab": lambda: scenes.slab_scene(64),
    "glass": lambda: scenes.small_scene("glass", 64),
    "mismatch": lambda: scenes.small_scene("mismatch", 64),
    "heightfield": lambda: scenes.small_scene("heightfield", 64),
}
# the fields a comparison needs, at the precision it needs
FIXTURE_DTYPE = np.dtype([("type", "i1"), ("layer", "i1"), ("aux", "<i4"), ("draws", "<u4"),
                          ("pos", "<f8", (3,)), ("dir", "<f8", (3,)), ("weight", "<f8")])


def tape_trace(inp, photon):
    return ol.trace(ol.ref_tape().xref_trace, inp, SEED, photon, 200000)


def main():
    out = dict(seed=SEED, photons=PHOTONS)
    for name, make in SCENES.items():
        inp = InputHandle(make())
        for p in range(PHOTONS):
            t = tape_trace(inp, p)
            f = np.zeros(len(t), FIXTURE_DTYPE)
            for k in FIXTURE_DTYPE.names:
                f[k] = t[k]
            out[f"{name}_{p}"] = f
    np.savez_compressed(Path(__file__).parent / "tape_traces.npz", **out)


if __name__ == "__main__":
    main()
