"""Train the layered-tissue surrogate on data the batched GPU Monte Carlo
generates, judge it against high-photon Monte Carlo on held-out
configurations, and write the model (with that judgement set) to one file.

    python -m xmcml_amd.train_surrogate --out xmcml_amd/data/surrogate_layered.npz

Training configurations and held-out configurations are drawn from
different seeds of the batched workload's distribution (2-5 layers, log-U mua
in [0.001, 1]/mm, log-U mus' in [0.5, 5]/mm, U g in [0.7, 0.95], U n in
[1.33, 1.5], log-U thickness in [0.05, 10] mm).  --data-out/--data-in keep
and reuse the simulated data so that training can be repeated without
simulating again.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import surrogate as S


def _configs_array(configs):
    a = np.zeros((len(configs), S.MAX_LAYERS, 5))
    n = np.zeros(len(configs), np.int32)
    for i, c in enumerate(configs):
        n[i] = len(c)
        a[i, :len(c)] = np.asarray(c)
    return a, n


def configs_from_array(a, n):
    return [[tuple(float(v) for v in a[i, j]) for j in range(int(n[i]))] for i in range(len(n))]


def evaluate(sur: S.Surrogate, X, Wref, Kref, min_count: int):
    mu, sig = sur.predict(X)
    valid = Kref[:, :S.RINGS] >= min_count
    err = S.ring_relative_error(np.exp(mu[:, :S.RINGS]), Wref[:, :S.RINGS], valid)
    tv = Kref[:, S.RINGS:] >= min_count
    terr = np.abs(np.exp(mu[:, S.RINGS:])[tv] - Wref[:, S.RINGS:][tv]) / Wref[:, S.RINGS:][tv]
    return {
        "ring_median_rel_err": float(np.median(err)),
        "ring_p90_rel_err": float(np.quantile(err, 0.9)),
        "ring_entries": int(valid.sum()),
        "tof_median_rel_err": float(np.median(terr)) if terr.size else None,
        "uncertainty": sur.uncertainty(sig),
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-configs", type=int, default=65536)
    ap.add_argument("--train-photons", type=int, default=30000)
    ap.add_argument("--eval-configs", type=int, default=512)
    ap.add_argument("--eval-photons", type=int, default=10_000_000)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=1e-5)
    ap.add_argument("--weight-decay", type=float, default=0.0)
    ap.add_argument("--hidden", type=int, default=512)
    ap.add_argument("--depth", type=int, default=4)
    ap.add_argument("--min-count", type=int, default=1000,
                    help="reference packets a ring needs to enter the error median")
    ap.add_argument("--flag-quantile", type=float, default=0.98,
                    help="uncertainty quantile over the training inputs above which inputs fall back to MC")
    ap.add_argument("--train-seed", type=int, default=1001)
    ap.add_argument("--eval-seed", type=int, default=2002)
    ap.add_argument("--data-in", type=Path, nargs="+",
                    help="simulated data sets to train on (concatenated); the judgement set is the first one's")
    ap.add_argument("--data-out", type=Path)
    ap.add_argument("--data-only", action="store_true", help="simulate and write --data-out, do not train")
    ap.add_argument("--eval-from", type=Path,
                    help="take the held-out configurations and their reference MC (layers, Wref, Kref) from this "
                         "model or data file instead of simulating them (also with --data-in)")
    ap.add_argument("--out", type=Path)
    ap.add_argument("--log-every", type=int, default=1000)
    args = ap.parse_args(argv)

    def progress(tag):
        t0 = time.perf_counter()

        def f(done, total):
            print(f"{tag}: {done}/{total} configurations, {time.perf_counter() - t0:.1f} s", flush=True)
        return f

    if not args.data_only and args.out is None:
        ap.error("--out is required unless --data-only")
    timings = {}
    if args.data_in:
        parts = [np.load(f, allow_pickle=False) for f in args.data_in]
        X = np.concatenate([z["X"] for z in parts])
        Y = np.concatenate([z["Y"] for z in parts])
        M = np.concatenate([z["M"] for z in parts]).astype(np.float32)
        for q in parts:
            for k, v in json.loads(bytes(q["timings"]).decode()).items():
                timings[k] = timings.get(k, 0.0) + v
        photons = {int(q["train_photons"]) for q in parts}
        if len(photons) != 1:
            ap.error("the data sets were simulated with different photon budgets")
        train_photons = photons.pop()
        if args.eval_from:
            # the judgement set of an earlier model or data file; plain MC at this training budget
            # on it is simulated here when the file does not hold it
            z = np.load(args.eval_from, allow_pickle=False)
            eval_a, eval_n, Wref, Kref = z["eval_layers"], z["eval_count"], z["Wref"], z["Kref"]
            if "W_lo" in z.files and int(z["train_photons"]) == train_photons:
                W_lo = z["W_lo"]
            else:
                t0 = time.perf_counter()
                _, _, _, W_lo = S.simulate(configs_from_array(eval_a, eval_n), train_photons,
                                           seed=args.eval_seed * 104729)
                timings["eval_mc_s"] = timings.get("eval_mc_s", 0.0) + time.perf_counter() - t0
        else:
            z = next(z for z in parts if z["eval_count"].size > 0)
            eval_a, eval_n, Wref, Kref = z["eval_layers"], z["eval_count"], z["Wref"], z["Kref"]
            W_lo = z["W_lo"]
    else:
        train_cfg = S.draw_configs(args.train_configs, args.train_seed)
        eval_cfg = S.draw_configs(args.eval_configs, args.eval_seed)
        given = None
        if args.eval_from:
            given = np.load(args.eval_from, allow_pickle=False)
            eval_cfg = configs_from_array(given["eval_layers"], given["eval_count"])
        t0 = time.perf_counter()
        Y, M, _, _ = S.simulate(train_cfg, args.train_photons, seed=args.train_seed * 7919,
                                progress=progress("training data"))
        timings["train_mc_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        Kref = Wref = W_lo = np.zeros((0, S.N_OUTPUTS))
        if eval_cfg:
            if given is not None:
                Kref, Wref = given["Kref"], given["Wref"]
            else:
                _, _, Kref, Wref = S.simulate(eval_cfg, args.eval_photons, seed=args.eval_seed * 7919, chunk=128,
                                              progress=progress("held-out reference"))
            # plain MC at the training budget on the held-out configurations
            _, _, _, W_lo = S.simulate(eval_cfg, args.train_photons, seed=args.eval_seed * 104729)
        timings["eval_mc_s"] = time.perf_counter() - t0
        X = S.encode_many(train_cfg)
        eval_a, eval_n = _configs_array(eval_cfg)
        train_photons = args.train_photons
        if args.data_out:
            args.data_out.parent.mkdir(parents=True, exist_ok=True)
            np.savez_compressed(args.data_out, X=X, Y=Y, M=M.astype(np.uint8), eval_layers=eval_a,
                                eval_count=eval_n, Wref=Wref, Kref=Kref, W_lo=W_lo,
                                train_photons=np.int64(args.train_photons),
                                timings=np.frombuffer(json.dumps(timings).encode(), np.uint8))
        if args.data_only:
            print(json.dumps({"configs": len(X), "photons": args.train_photons, **timings}), flush=True)
            return 0
    eval_cfg = configs_from_array(eval_a, eval_n)
    Xe = S.encode_many(eval_cfg)

    sur, log = S.Surrogate.fit(X, Y, M, steps=args.steps, batch=args.batch, lr=args.lr, lr_final=args.lr_final,
                               weight_decay=args.weight_decay, hidden=args.hidden, depth=args.depth,
                               log_every=args.log_every)
    _, sig_train = sur.predict(X[:min(len(X), 65536)])
    sur.sigma_max = float(np.quantile(sur.uncertainty(sig_train), args.flag_quantile))
    ev = evaluate(sur, Xe, Wref, Kref, args.min_count)
    unc = ev.pop("uncertainty")
    flagged = unc > sur.sigma_max
    valid = Kref[:, :S.RINGS] >= args.min_count
    mc_err = S.ring_relative_error(W_lo[:, :S.RINGS], Wref[:, :S.RINGS], valid & (W_lo[:, :S.RINGS] > 0))
    summary = {
        "train_configs": int(len(X)), "train_photons": train_photons,
        "eval_configs": int(len(eval_cfg)), "eval_photons": args.eval_photons, "min_count": args.min_count,
        "steps": log.steps, "batch": args.batch, "fit_seconds": log.seconds,
        "loss_first": log.first_loss, "loss_last": log.last_loss,
        **ev,
        "ring_median_rel_err_unflagged": float(np.median(S.ring_relative_error(
            np.exp(sur.predict(Xe[~flagged])[0][:, :S.RINGS]), Wref[~flagged, :S.RINGS], valid[~flagged])))
        if (~flagged).any() else None,
        "flagged_fraction": float(flagged.mean()), "sigma_max": sur.sigma_max,
        "plain_mc_at_train_budget_ring_median_rel_err": float(np.median(mc_err)),
        **timings,
    }
    args.out.parent.mkdir(parents=True, exist_ok=True)
    sur.save(args.out, meta=summary)
    # the judgement set travels with the model
    z = dict(np.load(args.out, allow_pickle=False))
    z.update(eval_layers=eval_a, eval_count=eval_n, Wref=Wref, Kref=Kref)
    np.savez_compressed(args.out, **z)
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
