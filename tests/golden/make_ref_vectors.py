"""Writes tests/golden/ref_vectors.npz from the reference build oracle/_ref
(the reference's own sources compiled by oracle/Makefile): tallies (as sha256 digests where large) of
300 photons of the 3-layer slab (termination KILL, per-photon streams, seed 777),
per-step traces of its first 3 photons, and the BVH arrays of the small
heightfield.  Run in a container that has the reference tree:

    make -C oracle ref && python tests/golden/make_ref_vectors.py
"""
import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT), str(ROOT / "tests")]

import oracle_lib as ol  # noqa: E402
from xmcml_amd import scenes  # noqa: E402
from xmcml_amd.abi import InputHandle, OutputHandle, TERMINATE_KILL  # noqa: E402


def digest(a) -> np.ndarray:
    """sha256 of the array's little-endian bytes: an exact check without the array."""
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), dtype=np.uint8)


def main():
    n, seed, n_traces = 300, 777, 3
    spec = scenes.slab_scene(n)
    spec.termination = TERMINATE_KILL
    inp = InputHandle(spec)
    o = OutputHandle(spec, scattering=True, trajectory=True)
    assert ol.ref().xref_run(inp.ptr, o.ptr, seed, 0, n, 0) == 0
    out = dict(n=n, seed=seed, n_traces=n_traces, absorption=digest(o.absorption), scattering=digest(o.scattering),
               trajectory=digest(o.trajectory), absorption_sum=o.absorption.sum(),
               det_weight=o.det_weight(), bin_weight=o.bin_weight(), det_count=o.det_count())
    for p in range(n_traces):
        out[f"trace{p}"] = ol.trace(ol.ref().xref_trace, inp, seed, p, 50000)
    hs = scenes.small_scene("heightfield", 10)
    A = ol.bvh_arrays(ol.ref().xref_bvh, InputHandle(hs))
    for k in ("boxes", "links", "tsurf", "tidx"):
        out["bvh_" + k] = digest(A[k])
    np.savez_compressed(Path(__file__).parent / "ref_vectors.npz", **out)


if __name__ == "__main__":
    main()
