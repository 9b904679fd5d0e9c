"""Learned warm start, end to end: training pairs from the pipeline, GPU training, held-out evaluation.

    python -m pct_reconstruction_amd.warmstart --train 400 --test 20 --histories 10000000 \
        --weights warmstart.bin --report warmstart.json

Workload (learned-128): random ellipsoid phantoms (6-14 ellipsoids, RSP 0-1.8) on 128x128x64 at
2 mm, 10M seeded histories each, MSC hull, FBP x0.  Training pair = (FBP image, hull, DROP image
after 40 iterations from the FBP start).  Evaluation per held-out seed:
  FBP start     x0 = FBP, `ref_iterations` DROP iterations -> residual r_ref = ||Ax - b||, time t_fbp;
  learned start x0 = net(FBP, hull) if it lowers ||Ax - b|| (else FBP), then DROP iterations until the
                residual is <= r_ref (at least one) -> iterations k, time t_net (net inference and the
                residual judgement included).
Reported: k, t_fbp / t_net, the relative L2 difference between the two final images.

With --input-iterations m > 0 the net's input is the iterate after m DROP iterations from the FBP
start instead of the FBP image (training pair = (x_m, hull, DROP image after 40 iterations)); the
learned start then runs m iterations, applies the net to x_m (kept only if it lowers ||Ax - b||) and
continues; its iteration count and time include the m iterations.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import configs, synth
from .learned import WarmStartNet
from .recon import Reconstruction


HEAD_SEED = 2024  # data seed of the head-50M benchmark


def _config(name: str, seed: int, n: int):
    """Parameters, phantom, beam of phantom `seed`; seed < 0 is the head-50M head phantom itself."""
    if name == "head-50M":
        if seed < 0:
            return configs.head_50m(n)
        return configs.learned_head50m(seed, n)
    return configs.learned_128(seed, n)


def _prepare(seed: int, n: int, config: str = "learned-128"):
    P, ph, bm, n = _config(config, seed, n)
    d = synth.generate(ph, bm, HEAD_SEED if seed < 0 else seed, n)
    r = Reconstruction(P)
    r.load_histories(d["t"], d["v"], d["u"], d["wepl"], d["angle_deg"])
    del d
    r.preprocessing()
    r.reconstruction_cuts()
    r.build_mlp_lists()
    r.synchronize()
    return P, r


def training_pair(seed: int, n: int, iterations: int = 40, input_iterations: int = 0, config: str = "learned-128"):
    """(net input, hull, DROP image after `iterations` from the FBP start) of phantom `seed`; the input
    is the FBP image, or the iterate after `input_iterations` DROP iterations."""
    P, r = _prepare(seed, n, config)
    try:
        hull = r.image("hull").astype(np.uint8)
        if input_iterations > 0:
            r.DROP(input_iterations)
            inp = r.image("x")
            r.DROP(iterations - input_iterations)
        else:
            inp = r.image("fbp")
            r.DROP(iterations)
        target = r.image("x")
    finally:
        r.close()
    return inp, hull, target


def _learned_start(net: WarmStartNet, r, m: int) -> dict:
    if m > 0:
        r.DROP(m)
        j = net.apply_iterate(r)
        return {"used_net": j["used_net"], "residual_net": j["residual_net"], "residual_fbp": j["residual_x"]}
    return net.apply(r)


def evaluate(net: WarmStartNet, seed: int, n: int, ref_iterations: int, long_iterations: int = 40,
             input_iterations: int = 0, config: str = "learned-128") -> dict:
    P, r = _prepare(seed, n, config)
    try:
        hull = r.image("hull").astype(bool)
        # reference image: `long_iterations` DROP iterations from the FBP start
        r.reset_iterate()
        r.DROP(long_iterations)
        x_long = r.image("x")
        # FBP start
        r.reset_iterate()
        r.synchronize()
        t0 = time.perf_counter()
        r.DROP(ref_iterations)
        r.synchronize()
        t_fbp = time.perf_counter() - t0
        r_ref = r.residual()
        x_fbp = r.image("x")
        r.reset_iterate()
        r0_fbp = r.residual()
        # learned start: count the iterations needed (residual checked after each)
        judge = _learned_start(net, r, input_iterations)
        k = input_iterations
        res = judge["residual_net"] if judge["used_net"] else judge["residual_fbp"]
        while k < input_iterations + 1 or (res > r_ref and k < 4 * ref_iterations):
            r.DROP(1)
            k += 1
            res = r.residual()
        x_net = r.image("x")
        # timed re-run: inference + residual judgement + k iterations
        r.reset_iterate()
        r.synchronize()
        t0 = time.perf_counter()
        _learned_start(net, r, input_iterations)
        r.DROP(k - input_iterations)
        r.synchronize()
        t_net = time.perf_counter() - t0
    finally:
        r.close()

    def rel(a, b):  # relative L2 difference inside the hull (DROP never changes voxels outside it)
        return float(np.linalg.norm((a - b)[hull]) / np.linalg.norm(b[hull]))

    return {"seed": seed, "used_net": judge["used_net"], "residual_x0_fbp": r0_fbp,
            "residual_x0_net": judge["residual_net"], "residual_target": r_ref, "residual_reached": res,
            "iterations_fbp": ref_iterations, "iterations_net": k, "t_fbp_s": t_fbp, "t_net_s": t_net,
            "speedup": t_fbp / t_net if t_net > 0 else None,
            "hull_rel_diff_net_vs_fbp": rel(x_net, x_fbp),
            f"hull_rel_err_fbp_vs_{long_iterations}it": rel(x_fbp, x_long),
            f"hull_rel_err_net_vs_{long_iterations}it": rel(x_net, x_long)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="learned-128", choices=["learned-128", "head-50M"],
                    help="learned-128 (128x128x64 @ 2 mm, 10M histories) or head-50M (256x256x96 @ 1 mm; "
                         "evaluated on the head-50M benchmark workload first, then on held-out random phantoms)")
    ap.add_argument("--train", type=int, default=400, help="training phantoms (seeds 0..train-1)")
    ap.add_argument("--test", type=int, default=20, help="held-out phantoms (seeds 100000..)")
    ap.add_argument("--histories", type=int, default=10_000_000)
    ap.add_argument("--target-iterations", type=int, default=40)
    ap.add_argument("--ref-iterations", type=int, default=20)
    ap.add_argument("--steps", type=int, default=4000, help="Adam steps (one phantom each)")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--weights", default=None, help="write (or with --eval-only read) the weights here")
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--input-iterations", type=int, default=1,
                    help="net input: FBP (0) or the iterate after this many DROP iterations")
    ap.add_argument("--report", default=None)
    args = ap.parse_args(argv)

    P, _, _, _ = _config(args.config, 0, 1)
    net = WarmStartNet(P.columns, P.rows, P.slices, seed=12345)
    report = {"workload": f"{args.config}: random ellipsoid phantoms, {P.columns}x{P.rows}x{P.slices}, "
                          f"{args.histories} histories each, FBP x0",
              "train": args.train, "test": args.test, "target_iterations": args.target_iterations,
              "ref_iterations": args.ref_iterations, "input_iterations": args.input_iterations}
    if args.eval_only:
        net.load(args.weights)
    else:
        t0 = time.perf_counter()
        for s in range(args.train):
            net.add_sample(*training_pair(s, args.histories, args.target_iterations, args.input_iterations,
                                          args.config))
            print(f"[warmstart] {s + 1}/{args.train} training pairs ({time.perf_counter() - t0:.0f} s)", flush=True)
        report["data_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        losses = []
        chunk = 250
        for k in range(0, args.steps, chunk):
            losses.append(net.train(min(chunk, args.steps - k), lr=args.lr, seed=k))
            print(f"[warmstart] step {k + chunk}: mean loss {losses[-1]:.4e}", flush=True)
        report["train_s"] = time.perf_counter() - t0
        report["loss_curve"] = losses
        if args.weights:
            net.save(args.weights)
    evals = []
    seeds = [100000 + j for j in range(args.test)]
    if args.config == "head-50M":
        seeds = [-1] + seeds  # the head-50M benchmark workload itself first
    for seed in seeds:
        e = evaluate(net, seed, args.histories, args.ref_iterations, input_iterations=args.input_iterations,
                     config=args.config)
        evals.append(e)
        print(f"[warmstart] test seed {seed}: {json.dumps(e)}", flush=True)
    sp = [e["speedup"] for e in evals if e["speedup"]]
    report["evaluations"] = evals
    report["median_speedup"] = float(np.median(sp)) if sp else None
    report["median_iterations_net"] = float(np.median([e["iterations_net"] for e in evals])) if evals else None
    report["net_used"] = sum(e["used_net"] for e in evals)
    net.close()
    out = json.dumps(report)
    if args.report:
        with open(args.report, "w") as f:
            f.write(out + "\n")
    print(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
