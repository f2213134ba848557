"""Photon-transport throughput on one MI355X, next to the CPU path on the
host's cores.

Workload (one step): `--photons` photons (default 1e8, the whole configuration, in one
persistent launch) of the 3-D surface configuration --
three conformal seeded heightfield boundaries triangulated on a 128 x 128 grid
over 40 x 40 mm plus a flat bottom (98,306 triangles in the BVH), the
3-layer tissue of the slab configuration, pencil beam, roulette at 1e-4 with
survival 0.1, 20 ring detectors of 0.5 mm, a 256-bin time-of-flight
histogram over 0-2 ns per ring.  Geometry, optics and tallies are resident in
HBM before the timed region; photon streams continue from step to step.  The
three heightfields share one uniform lattice and the bottom is a rectangle, so
each step's nearest surface comes from the engine's surface rule (the
triangles of the lattice cells the step covers, tested as the BVH leaves test
them); `--no-lattice-rule` times the BVH-walk path of the same workload.

Output: one JSON line.  `value` is photons/s over the timed steps;
`roofline` prices the transport kernel's algorithmic FP32 work (80 FLOP per
scatter event, 90 per boundary interaction, counted in-kernel) against the
FP32 peak, with the kernel time from HIP events around each launch;
`cpu_baseline` is the double-precision oracle (a restatement of the
reference's OpenMP LaunchOMP path) on a bounded sample of the same workload,
with the reference's own LaunchOMP rate derived from it through the same-core
ratio in oracle/reference_ratio.json.  After the timed steps (and the
surrogate's part), the default run also measures the other configurations --
one 1e8-photon launch of the slab with its absorption grid, the 1024 batched
configurations in their graph, one 1e8-photon launch of the 1.64M-triangle
head -- into `other_workloads` (`--no-other-workloads` skips them).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

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


WORKLOADS = {
    "heightfield": "3-D surface MC: 3 conformal heightfield boundaries (128x128 grid, 40x40 mm) + flat bottom, "
                   "98,306 triangles; 3 layers; pencil beam; roulette 1e-4/0.1; 20 rings + 256-bin TOF 0-2 ns",
    "head": "heavy geometry: 5 nested ellipsoids (90/75/65 mm outer, shells 5/7/2/3 mm), icosphere level 7 with "
            "1 mm seeded radial noise, 1,638,400 triangles; scalp/skull/CSF/grey/white; pencil beam; roulette "
            "1e-4/0.1; rings at 10-40 mm + 256-bin TOF 0-5 ns",
    "slab": "planar 3-layer slab (0.1/2/20 mm), pencil beam, roulette 1e-4/0.1, 20 rings of 0.5 mm, 100^3 "
            "absorption grid over 20x20x20 mm, 50-bin TOF",
    "batch": "1024 layered configurations of 2-5 layers (parameters and log-uniform photon budgets in [1e5, 1e7] "
             "from seed 3), 32 rings + 64 TOF bins each, one persistent launch replayed as a HIP graph",
}


def bench_batch(args) -> int:
    """The batched configurations: every budget in one persistent launch (a
    captured HIP graph); value = photons/s of the launch, plus configs/s."""
    print(json.dumps(measure_batch(args, args.cpu_seconds)), flush=True)
    return 0


def measure_batch(args, cpu_seconds: float) -> dict:
    from xmcml_amd import engine, scenes
    import oracle_lib as ol
    from xmcml_amd.abi import InputHandle, OutputHandle

    specs = scenes.batched_configs(1024, photon_scale=args.batch_scale)
    total = sum(s.photons for s in specs)
    print(f"bench: batch of {len(specs)} configurations, {total:.3e} photons", file=sys.stderr, flush=True)
    # launch shape from the autotuner: candidates timed on the same configurations at 5 % of their
    # budgets (cached per device and workload), the winner applied to the full graph
    b = engine.Batch(specs, tune=True, tune_scale=0.05)
    print(f"bench: launch shape {b.tuning['block']} threads x {b.tuning['blocks_per_cu']} blocks/CU "
          f"({'cached' if b.tuning['cached'] else 'tuned'})", file=sys.stderr, flush=True)
    b.run(args.seed)  # warm-up launch (and graph capture); the timed launch replays the same graph
    b.collect()
    b.stats()
    t0 = time.perf_counter()
    b.run(args.seed)
    outs = b.collect()
    elapsed = time.perf_counter() - t0
    st = b.stats()
    launched = b.get_launch()
    b.close()
    value = total / elapsed
    line = {
        "metric": "photons/s (batched configurations)", "value": value, "unit": "photons/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": elapsed * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 transport, fp64 tallies", "data": "synthetic",
        "config": {"workload": WORKLOADS["batch"], "configurations": len(specs), "photons": total},
        "configs_per_s": len(specs) / elapsed, "kernel_ms": st["kernelMs"],
        "launch_shape": {"block": b.tuning["block"], "blocks_per_cu": b.tuning["blocks_per_cu"],
                         "launched": list(launched), "cached": b.tuning["cached"],
                         "candidates_ms_at_5pct_budgets": b.tuning.get("timings")},
        "scatter_events_per_photon": st["scatterEvents"] / total,
    }
    if cpu_seconds > 0:
        # the double oracle on a sample: 32 configurations, 20,000 photons each, all host threads
        threads = host_threads()
        photons = seconds = 0.0
        t_end = time.perf_counter() + cpu_seconds
        k = 0
        while time.perf_counter() < t_end and k < 256:
            sp = specs[(k * 37) % len(specs)]
            o = OutputHandle(sp, absorption=False)
            c = ol.run_oracle(sp, InputHandle(sp), o, args.seed, 0, 20000, 64, threads)
            photons += c.photons
            seconds += c.seconds
            k += 1
        line["cpu_baseline"] = {"value": photons / seconds, "unit": "photons/s", "cores": threads,
                                "nproc": os.cpu_count(), "cpu_model": cpu_model(), "kind": "port",
                                "sample": f"{k} configurations of the batch x 20,000 photons, double precision"}
        line["speedup_vs_cpu"] = value / line["cpu_baseline"]["value"]
    return line


def measure_scene(name: str, photons: int, seed: int) -> dict:
    """One untimed launch of a tenth of `photons`, then one launch of `photons`
    of a single-scene workload: photons/s, kernel time, FP32 roofline."""
    from xmcml_amd import engine, scenes
    make = {"head": scenes.head_scene, "slab": scenes.slab_scene}[name]
    spec = make(photons)
    t_build = time.perf_counter()
    sc = engine.Scene(spec, absorption=name == "slab")
    t_build = time.perf_counter() - t_build
    sc.run(seed, 0, photons // 10)
    sc.sync()
    sc.collect()
    sc.stats()
    t0 = time.perf_counter()
    sc.run(seed, photons, photons)
    sc.sync()
    elapsed = time.perf_counter() - t0
    st = sc.stats()
    sc.collect()
    sc.close()
    flop = FLOP_SCATTER * st["scatterEvents"] + FLOP_BOUNDARY * st["boundaryEvents"]
    achieved = flop / (st["kernelMs"] / 1e3) / 1e12
    return {"workload": WORKLOADS[name], "value": photons / elapsed, "unit": "photons/s", "photons": photons,
            "ms_per_step": elapsed * 1e3, "kernel_ms": st["kernelMs"], "scene_setup_s": t_build,
            "roofline_achieved_tflops": achieved, "roofline_frac": achieved / PEAK_FP32_TFLOPS,
            "scatter_events_per_photon": st["scatterEvents"] / photons}


def host_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def reference_equivalent(value: float, workload: str):
    """The reference's own LaunchOMP is slower than the restatement timed
    here; oracle/reference_ratio.json (oracle/measure_reference_ratio.py, run
    where the reference builds) holds the same-core ratio.  Returns the
    restatement's rate divided by it, with the ratio's provenance."""
    f = ROOT / "oracle" / "reference_ratio.json"
    try:
        r = json.loads(f.read_text())
        w = r["workloads"][workload]
    except (OSError, KeyError, ValueError):
        return None
    return {"value": value / w["port_over_reference"], "port_over_reference": w["port_over_reference"],
            "measured_on": f"{r['threads']} threads of {r['cpu_model']}: restatement {w['port_photons_per_s']:.4g} "
                           f"vs the reference's LaunchOMP compiled from its sources {w['reference_photons_per_s']:.4g} "
                           f"photons/s ({r['termination']})"}


def cpu_baseline(spec, target_seconds: float, seed: int, absorption: bool = False, workload: str = "") -> dict:
    """The double-precision oracle on all host threads, on a bounded sample."""
    import oracle_lib as ol
    from xmcml_amd.abi import InputHandle, OutputHandle

    threads = host_threads()
    inp = InputHandle(spec)
    n = max(200, 20 * threads)
    out = OutputHandle(spec, absorption=absorption)
    c = ol.run_oracle(spec, inp, out, seed, 0, n, 64, threads)  # warm-up, not counted
    rate = c.photons / max(c.seconds, 1e-9)
    first, photons, seconds, scatter = n, 0, 0.0, 0
    out = OutputHandle(spec, absorption=absorption)
    while seconds < target_seconds:
        chunk = int(max(n, rate * (target_seconds - seconds)))
        c = ol.run_oracle(spec, inp, out, seed, first + photons, chunk, 64, threads)
        photons += c.photons
        seconds += c.seconds
        scatter += c.scatterEvents
        rate = photons / seconds
    value = photons / seconds
    return {
        "value": value,
        "unit": "photons/s",
        "cores": threads,
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model(),
        "kind": "port",
        "reference_equivalent": reference_equivalent(value, workload),
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


def surrogate_line(model_path: Path, queries: int, mc_configs: int, mc_photons: int, answer_queries: int,
                   fallback_photons: int) -> dict:
    """The learned half of the metric: configurations/s of the surrogate's
    inference kernels (features resident on the device) and its median
    ring-reflectance error against 1e7-photon Monte Carlo on held-out
    configurations, next to plain batched Monte Carlo's configurations/s at
    mc_photons photons each, measured now on the same GPU."""
    from xmcml_amd import engine
    from xmcml_amd import surrogate as S
    from xmcml_amd.train_surrogate import configs_from_array, evaluate

    z = np.load(model_path, allow_pickle=False)
    sur = S.Surrogate.load(model_path, max_batch=65536)
    cfg = configs_from_array(z["eval_layers"], z["eval_count"])
    Xe = S.encode_many(cfg)
    ev = evaluate(sur, Xe, z["Wref"], z["Kref"], int(sur.meta.get("min_count", 1000)))
    unc = ev.pop("uncertainty")
    flagged = unc > sur.sigma_max
    mu, _ = sur.predict(Xe)
    # inference throughput on a resident batch of `queries` feature rows
    Xq = np.tile(Xe, (queries // len(Xe) + 1, 1))[:queries]
    xq = engine.DeviceArray(Xq.shape, Xq)
    out = sur.predict_device(xq)
    engine.device_sync()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        sur.predict_device(xq, out)
    engine.device_sync()
    t_inf = (time.perf_counter() - t0) / reps
    # the same forward as one GEMM launch per layer (activations through HBM)
    sur.mlp.forward(xq, out, per_layer=True)
    engine.device_sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        sur.mlp.forward(xq, out, per_layer=True)
    engine.device_sync()
    t_layers = (time.perf_counter() - t0) / reps
    # end to end through Surrogate.answer: `answer_queries` configurations of the batched
    # workload's distribution (one in 1000 pushed out of the sampling box), host arrays in,
    # answers out -- vectorised encoding, copies, inference, the uncertainty check, and plain
    # batched MC for every flagged or out-of-range configuration
    La, ca = S.draw_config_arrays(answer_queries, 4242)
    La[::1000, 0, 1] = 2.0  # mua 2/mm: outside the box
    t0 = time.perf_counter()
    _, _, mc_mask, stages = sur.answer_arrays(La, ca, mc_photons=fallback_photons, seed=77)
    t_e2e = time.perf_counter() - t0
    # plain batched Monte Carlo on held-out configurations
    t0 = time.perf_counter()
    S.simulate(cfg[:mc_configs], mc_photons, seed=99, chunk=mc_configs)
    t_mc = time.perf_counter() - t0
    mc_rate = mc_configs / t_mc
    inf_rate = queries / t_inf
    h, d = sur.mlp.hidden, sur.mlp.depth
    fl = 2.0 * (S.N_INPUTS * h + (d - 1) * h * h + h * 2 * S.N_OUTPUTS)  # algorithmic FLOP per configuration
    return {
        "configs_per_s": inf_rate,
        "unit": "configs/s",
        "median_ring_rel_err": ev["ring_median_rel_err"],
        "p90_ring_rel_err": ev["ring_p90_rel_err"],
        "tof_median_rel_err": ev["tof_median_rel_err"],
        "judged_against": f"{len(cfg)} held-out configurations, {sur.meta.get('eval_photons', 10_000_000):.0e}-photon "
                          f"MC, rings with >= {sur.meta.get('min_count', 1000)} detected packets "
                          f"({ev['ring_entries']} entries)",
        "flagged_for_mc_fraction": float(flagged.mean()),
        "median_ring_rel_err_unflagged": float(np.median(S.ring_relative_error(
            np.exp(mu[~flagged, :S.RINGS]), z["Wref"][~flagged, :S.RINGS],
            z["Kref"][~flagged, :S.RINGS] >= int(sur.meta.get("min_count", 1000))))) if (~flagged).any() else None,
        "queries_resident": queries,
        "inference_ms": t_inf * 1e3,
        "inference_tflops": fl * queries / t_inf / 1e12,
        "inference_frac_of_bf16_peak": fl * queries / t_inf / 2.5e15,
        "inference_kernel": "fused: whole network per 64-row tile, activations in LDS",
        "inference_ms_per_layer_launches": t_layers * 1e3,
        "plain_mc_configs_per_s": mc_rate,
        "plain_mc_photons_per_config": mc_photons,
        "speedup_vs_plain_mc": inf_rate / mc_rate,
        "answer": {
            "queries": answer_queries,
            "configs_per_s": answer_queries / t_e2e,
            "speedup_vs_plain_mc": answer_queries / t_e2e / mc_rate,
            "answered_by_mc": int(mc_mask.sum()),
            "fallback_photons_per_config": fallback_photons,
            "seconds": {"total": t_e2e, **stages},
            "configs_per_s_without_fallback": answer_queries / max(t_e2e - stages["fallback_s"], 1e-9),
            "note": "end to end from host arrays, fallback Monte Carlo included; with a fraction F of "
                    "queries answered by MC the ratio to plain MC is bounded by 1/F",
        },
        "trained_on": f"{sur.meta.get('train_configs')} configurations x {sur.meta.get('train_photons')} photons "
                      f"of batched GPU MC",
    }


def traffic_from_profiles(photons_per_launch: int, workload: str):
    """HBM bytes per transport launch: the committed rocprofv3 FETCH_SIZE +
    WRITE_SIZE summary (per photon, copy-calibrated) times the photons of one
    launch; None for workloads without a profile."""
    f = ROOT / "profiles" / f"transport_{workload}_pmc.json"
    if f.exists():
        try:
            return json.loads(f.read_text())["hbm_bytes_per_photon"] * photons_per_launch
        except Exception:
            return None
    return None


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=int, default=None,
                    help="photons per step, one launch (default: 1e8 -- the heightfield configuration's count, "
                         "and a tenth of the head configuration's 1e9)")
    ap.add_argument("--workload", choices=["heightfield", "head", "slab", "batch"], default="heightfield",
                    help="heightfield: the 3-D surface configuration (the metric's workload); head: five nested "
                         "noisy ellipsoids of 327,680 triangles each; slab: the planar 3-layer configuration with "
                         "its 100^3 absorption grid; batch: the 1024 layered configurations in one launch")
    ap.add_argument("--batch-scale", type=float, default=1.0, help="batch workload: photon budgets times this")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample length (0 = skip)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--no-slab-grid", action="store_true",
                    help="build no separating slabs near the mesh (A/B of the walk skipping; same results)")
    ap.add_argument("--no-other-workloads", action="store_true",
                    help="skip the other configurations' lines (slab, batch, head), measured after the "
                         "default workload's timed steps")
    ap.add_argument("--no-lattice-rule", action="store_true",
                    help="walk the BVH on the heightfield and slab workloads instead of testing the triangles of "
                         "the lattice cells each step covers (the interleaved walk kernel; same hits)")
    ap.add_argument("--seed", type=int, default=777)
    ap.add_argument("--surrogate-model", type=Path, default=ROOT / "xmcml_amd" / "data" / "surrogate_layered.npz")
    ap.add_argument("--surrogate-queries", type=int, default=1 << 20)
    ap.add_argument("--mc-configs", type=int, default=64)
    ap.add_argument("--mc-photons", type=int, default=1_000_000)
    ap.add_argument("--answer-queries", type=int, default=1_000_000,
                    help="configurations answered end to end through Surrogate.answer (fallback MC included)")
    ap.add_argument("--fallback-photons", type=int, default=100_000,
                    help="photons per configuration of the fallback MC in the answer measurement (1e5: plain MC "
                         "there has 2.7 %% median ring error; the network's training targets used 1e6)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the steps, run one device-to-device copy of this many MiB (a known-size "
                         "kernel that scales the HBM counters of a profiler pass)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        return 2

    from xmcml_amd import engine, scenes

    if args.photons is None:
        # a launch ends with its slowest last photons: a tail that is a smaller share of a
        # longer launch (1e8 photons = 381 per lane of the persistent grid)
        args.photons = 100_000_000
    if args.workload == "batch":
        return bench_batch(args)
    make = {"heightfield": scenes.heightfield_scene, "head": scenes.head_scene, "slab": scenes.slab_scene}[args.workload]
    absorption = args.workload == "slab"
    spec = make(args.photons)
    print(f"bench: building scene ({spec.num_triangles} triangles)", file=sys.stderr, flush=True)
    sc = engine.Scene(spec, absorption=absorption, block_size=args.block, slab_grid=not args.no_slab_grid,
                      lattice_rule=not args.no_lattice_rule)
    # the heightfield's surfaces are lattice heightfields and a rectangle, the slab's rectangles: both
    # run the plain kernel under the surface rule; the head (closed meshes), and the heightfield
    # without its lattice rule, run the interleaved walk kernel
    walking = args.workload == "head" or (args.workload == "heightfield" and args.no_lattice_rule)
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
        # FP32 vector-ALU issue bound (SQ_ACTIVE_INST_VALU saturates): no matrix work, and the HBM
        # traffic is latency that other lanes hide; priced against the published FP32 peak
        "bound": "valu",
        "compute": "fp32 vector ALU (no matrix work on this path)",
        "achieved": achieved,
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32_TFLOPS,
        "frac_of_nonpacked_peak": achieved / PEAK_FP32_NONPACKED_TFLOPS,
        "traffic": traffic_from_profiles(args.photons, args.workload + ("_walk" if walking and args.workload != "head"
                                                                         else "")),
        "kernel": "xmc::transport_kernel<false, false, true>" if walking else "xmc::transport_kernel<false, false, false>",
        "nearest_surface": "BVH walk (interleaved, walk-skipping grids)" if walking else "surface rule (lattice cells / "
                           "rectangles)",
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
            "workload": WORKLOADS[args.workload],
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
        cb = cpu_baseline(make(0), args.cpu_seconds, args.seed + 1, absorption, args.workload)
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu"] = value / cb["value"]
        if cb["reference_equivalent"]:
            line["speedup_vs_cpu_reference"] = value / cb["reference_equivalent"]["value"]
    else:
        line["cpu_baseline"] = None
    line["roofline"] = roofline
    if args.surrogate_model and args.surrogate_model.exists() and args.surrogate_queries > 0:
        print("bench: surrogate inference and plain batched MC", file=sys.stderr, flush=True)
        line["surrogate"] = surrogate_line(args.surrogate_model, args.surrogate_queries, args.mc_configs,
                                           args.mc_photons, args.answer_queries, args.fallback_photons)
    else:
        line["surrogate"] = None
    if args.workload == "heightfield" and not args.no_other_workloads:
        # the other configurations, each after the timed region above: one launch of the slab's
        # 1e8 photons with its absorption grid, the 1024 batched configurations in their graph,
        # and 1e8 photons of the 1.64M-triangle head
        other = {}
        print("bench: other workloads: slab", file=sys.stderr, flush=True)
        other["slab"] = measure_scene("slab", 100_000_000, args.seed)
        print("bench: other workloads: batch", file=sys.stderr, flush=True)
        b = measure_batch(args, 0.0)
        other["batch"] = {k: b[k] for k in ("value", "unit", "ms_per_step", "configs_per_s", "kernel_ms",
                                            "launch_shape", "scatter_events_per_photon")}
        other["batch"]["workload"] = WORKLOADS["batch"]
        print("bench: other workloads: head", file=sys.stderr, flush=True)
        other["head"] = measure_scene("head", 100_000_000, args.seed)
        line["other_workloads"] = other
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
