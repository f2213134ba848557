"""Photon-transport throughput on one MI355X, next to the CPU path on the
host's cores.

Workload (one step): `--photons` photons of the 3-D surface configuration --
three conformal seeded heightfield boundaries triangulated on a 128 x 128 grid
over 40 x 40 mm plus a flat bottom (98,306 triangles in the BVH), the
3-layer tissue of the slab configuration, pencil beam, roulette at 1e-4 with
survival 0.1, 20 ring detectors of 0.5 mm, a 256-bin time-of-flight
histogram over 0-2 ns per ring.  Ten steps are the 1e8 photons of that
configuration.  Geometry, optics and tallies are resident in HBM before the
timed region; photon streams continue from step to step.

Output: one JSON line.  `value` is photons/s over the timed steps;
`roofline` prices the transport kernel's algorithmic FP32 work (80 FLOP per
scatter event, 90 per boundary interaction, counted in-kernel) against the
FP32 peak, with the kernel time from HIP events around each launch;
`cpu_baseline` is the double-precision oracle (a restatement of the
reference's OpenMP LaunchOMP path) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

METRIC = "photons/s on 1 MI355X vs host CPU; surrogate configs/s at ≤2% err vs MC"
FLOP_SCATTER = 80.0
FLOP_BOUNDARY = 90.0
PEAK_FP32_TFLOPS = 157.3        # MI355X FP32 vector, packed-counted
PEAK_FP32_NONPACKED_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def host_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(spec, target_seconds: float, seed: int) -> dict:
    """The double-precision oracle on all host threads, on a bounded sample."""
    import oracle_lib as ol
    from xmcml_amd.abi import InputHandle, OutputHandle

    threads = host_threads()
    inp = InputHandle(spec)
    n = max(200, 20 * threads)
    out = OutputHandle(spec, absorption=False)
    c = ol.run_oracle(spec, inp, out, seed, 0, n, 64, threads)  # warm-up, not counted
    rate = c.photons / max(c.seconds, 1e-9)
    first, photons, seconds, scatter = n, 0, 0.0, 0
    out = OutputHandle(spec, absorption=False)
    while seconds < target_seconds:
        chunk = int(max(n, rate * (target_seconds - seconds)))
        c = ol.run_oracle(spec, inp, out, seed, first + photons, chunk, 64, threads)
        photons += c.photons
        seconds += c.seconds
        scatter += c.scatterEvents
        rate = photons / seconds
    return {
        "value": photons / seconds,
        "unit": "photons/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{photons} photons of the same workload (photons {first}..{first + photons - 1} of seed "
                  f"{seed}), double precision, OpenMP dynamic schedule, {seconds:.1f} s",
        "scatter_events_per_photon": scatter / photons,
        "reflectance": float(out.det_weight().sum() / photons),
    }


def copy_calibration(nbytes: int) -> None:
    """One device-to-device hipMemcpy of `nbytes` (the runtime's blit copy
    kernel): it reads and writes exactly nbytes, which scales the HBM counters
    of a profiler pass over the same process."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for name, ptr in (("src", src), ("dst", dst)):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)) != 0:
            raise RuntimeError(f"hipMalloc of the calibration {name} buffer failed")
    try:
        if hip.hipMemset(src, 1, C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
        if hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)


def traffic_from_profiles():
    """HBM bytes per transport launch from the committed rocprofv3 PMC
    summary, when one exists."""
    f = ROOT / "profiles" / "transport_heightfield_pmc.json"
    if f.exists():
        try:
            return json.loads(f.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            return None
    return None


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=int, default=10_000_000, help="photons per step")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample length (0 = skip)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--seed", type=int, default=777)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the steps, run one device-to-device copy of this many MiB (a known-size "
                         "kernel that scales the HBM counters of a profiler pass)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        return 2

    from xmcml_amd import engine, scenes

    spec = scenes.heightfield_scene(args.photons)
    print(f"bench: building scene ({spec.num_triangles} triangles)", file=sys.stderr, flush=True)
    sc = engine.Scene(spec, absorption=False, block_size=args.block)
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib << 20)
        print(f"bench: calibration copy of {args.copy_calibration_mib} MiB done", file=sys.stderr, flush=True)
    first = 0
    for _ in range(args.warmup):
        sc.run(args.seed, first, args.photons)
        first += args.photons
    sc.sync()
    sc.collect()
    sc.stats()  # reset counters and kernel timers
    print(f"bench: warmup done, timing {args.steps} steps of {args.photons} photons", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    for _ in range(args.steps):
        sc.run(args.seed, first, args.photons)
        first += args.photons
    sc.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    st = sc.stats()
    out = sc.collect()
    photons = args.steps * args.photons
    value = photons / elapsed

    flop = FLOP_SCATTER * st["scatterEvents"] + FLOP_BOUNDARY * st["boundaryEvents"]
    launch_s = st["kernelMs"] / 1e3 / args.steps
    achieved = flop / args.steps / launch_s / 1e12
    roofline = {
        "bound": "mfma",
        "compute": "fp32 vector ALU (no matrix work on this path)",
        "achieved": achieved,
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32_TFLOPS,
        "frac_of_nonpacked_peak": achieved / PEAK_FP32_NONPACKED_TFLOPS,
        "traffic": traffic_from_profiles(),
        "kernel": "xmc::transport_kernel<false>",
        "kernel_ms_per_launch": launch_s * 1e3,
        "flop_per_launch": flop / args.steps,
        "scatter_events_per_photon": st["scatterEvents"] / photons,
        "boundary_events_per_photon": st["boundaryEvents"] / photons,
    }
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 transport, fp64 tallies",
        "data": "synthetic",
        "config": {
            "workload": "3-D surface MC: 3 conformal heightfield boundaries (128x128 grid, 40x40 mm) + flat "
                        "bottom, 98,306 triangles; 3 layers; pencil beam; roulette 1e-4/0.1; 20 rings + "
                        "256-bin TOF 0-2 ns",
            "photons_per_step": args.photons,
            "triangles": spec.num_triangles,
            "parallelism": "single GPU, persistent kernel",
        },
        "reflectance": float(out.det_weight().sum() / photons),
        "ledger_check": float(out.specular + (st["wAbsorbed"] + st["wEscaped"] + st["wLeftArea"] + st["wKilled"]
                                              + st["wRouletteLost"] + st["wOther"] - st["wRouletteGain"]) / photons),
    }
    sc.close()
    if args.cpu_seconds > 0:
        print("bench: timing the CPU baseline", file=sys.stderr, flush=True)
        cb = cpu_baseline(scenes.heightfield_scene(0), args.cpu_seconds, args.seed + 1)
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu"] = value / cb["value"]
    else:
        line["cpu_baseline"] = None
    line["roofline"] = roofline
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
