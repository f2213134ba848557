"""Measures how much faster the oracle's 64-bit restatement of LaunchOMP runs
than the reference's own LaunchOMP (base/xmcml/xmcmlLauncher/
launcher_omp.cpp:123-163, compiled unmodified into oracle/_ref by
oracle/Makefile) on the same inputs, the same threads and the same core, and
writes oracle/reference_ratio.json.  bench.py's CPU baseline times the
restatement on the GPU host (the reference build does not travel there) and
divides by this ratio to state the reference-equivalent rate.

TEST INFRASTRUCTURE: run in a container that has the reference tree,

    make -C oracle ref && python oracle/measure_reference_ratio.py

Both sides use xmcml's own termination (kill below minWeight 1e-4, the only
one LaunchOMP has); photons/s are compared, and the walks being the same
algorithm (bit-identical on the same streams), the events per photon agree
statistically.  The two sides alternate, each timed three times; the ratio is
of the medians.
"""
from __future__ import annotations

import json
import os
import platform
import statistics
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path[:0] = [str(ROOT), str(ROOT / "tests")]

import oracle_lib as ol  # noqa: E402
from xmcml_amd import scenes  # noqa: E402
from xmcml_amd.abi import InputHandle, OutputHandle, TERMINATE_KILL  # noqa: E402

WORKLOADS = {
    # the bench workloads, as bench.py's CPU baseline runs them
    "heightfield": (lambda n: scenes.heightfield_scene(n), False, 60_000),
    "slab": (lambda n: scenes.slab_scene(n), True, 300_000),
}


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main() -> int:
    threads = len(os.sched_getaffinity(0))
    out = {"threads": threads, "cpu_model": cpu_model(), "termination": "kill below 1e-4 (LaunchOMP's own)",
           "workloads": {}}
    for name, (make, absorption, n) in WORKLOADS.items():
        spec = make(n)
        spec.termination = TERMINATE_KILL
        inp = InputHandle(spec)
        port, ref = [], []
        for rep in range(3):
            o = OutputHandle(spec, absorption=absorption)
            c = ol.run_oracle(spec, inp, o, 100 + rep, 0, n, 64, threads)
            port.append(c.photons / c.seconds)
            o = OutputHandle(spec, absorption=absorption)
            inp.c.numberOfPhotons = n
            sec = ol.ref().xref_launch_omp(inp.ptr, o.ptr, 100 + rep, threads)
            ref.append(n / sec)
            print(f"{name} rep {rep}: port {port[-1]:.4g} photons/s, reference {ref[-1]:.4g}", flush=True)
        p, r = statistics.median(port), statistics.median(ref)
        out["workloads"][name] = {"photons": n, "absorption_grid": absorption, "port_photons_per_s": p,
                                  "reference_photons_per_s": r, "port_over_reference": p / r}
    (Path(__file__).parent / "reference_ratio.json").write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
