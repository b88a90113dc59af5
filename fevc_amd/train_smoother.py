"""Train the learned CMG smoother weights on the GPU and write them to
fevc_amd/smoother_params.json (built into the library by fevc_amd/build.py).

Training data are generated by the method itself: worker-firm panels of the
project's generator (200 panels of 50k-1M workers, T = 5, mobility
0.05-0.3, Pareto firm sizes), their hybrid Laplacians and CMG hierarchies built by the engine, and
error vectors that the plain V-cycle leaves behind.  Every step is one Adam
step of the device network on one hierarchy
(vckss_mi355x_smoother_train_v1).  Held-out panels then report the PCG
iterations of a 16-probe calibration batch with and without the weights.

    python -m fevc_amd.train_smoother --out smoother_params.json
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import fevc_amd as F  # noqa: E402
from fevc_amd import synth  # noqa: E402

COLS = ("worker", "firm", "deletion", "outcome", "frequency", "target")
FEATURES, HIDDEN, LAYERS, OMEGA_MAX = 8, 64, 3, 0.95
# parameter layout of the device network: W0[H][F], b0, then U_t, V_t, c_t of
# each message-passing layer, then a, a0
SIZES = ((("W0", HIDDEN * FEATURES), ("b0", HIDDEN))
         + tuple(x for t in range(1, LAYERS + 1)
                 for x in ((f"U{t}", HIDDEN * HIDDEN), (f"V{t}", HIDDEN * HIDDEN), (f"c{t}", HIDDEN)))
         + (("a", HIDDEN), ("a0", 1)))
COUNT = sum(s for _, s in SIZES)


def initial_params(seed: int, scale: float = 0.1) -> np.ndarray:
    """Small random weights; a = 0 and a0 = logit(0.5 / 0.95), so training
    starts from the plain cycle's 0.5."""
    rng = np.random.default_rng(seed)
    theta = np.zeros(COUNT)
    o = 0
    for name, size in SIZES:
        if name == "W0" or name[0] in "UV":
            fan_in = FEATURES if name == "W0" else HIDDEN
            theta[o:o + size] = rng.standard_normal(size) * scale / math.sqrt(fan_in)
        o += size
    q = 0.5 / OMEGA_MAX
    theta[-1] = math.log(q / (1.0 - q))
    return theta


def training_panels(count: int, seed: int):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        workers = int(math.exp(rng.uniform(math.log(50_000), math.log(1_000_000))))
        out.append(synth.PanelSpec(workers=workers, firms=max(workers // 20, 50), periods=5,
                                   mobility=float(rng.uniform(0.05, 0.3)),
                                   pareto_shape=float(rng.uniform(1.05, 1.3)), size_cap=5000.0,
                                   seed=10_000 + i, name=f"train-{i}"))
    return out


def held_out_panels():
    return [synth.PanelSpec(workers=400_000, firms=20_000, periods=5, mobility=0.12, pareto_shape=1.15,
                            seed=9001, name="held-out-400k"),
            synth.PanelSpec(workers=1_000_000, firms=50_000, periods=5, mobility=0.1, pareto_shape=1.1,
                            seed=9002, name="held-out-1M"),
            synth.PanelSpec(workers=200_000, firms=8_000, periods=5, mobility=0.25, pareto_shape=1.3,
                            seed=9003, name="held-out-200k")]


def engine(spec):
    d = synth.generate(spec)
    return F.Engine(*[d[k] for k in COLS])


def log(msg):
    print(f"[train {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--panels", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="Adam steps per panel visit")
    ap.add_argument("--count", type=int, default=0,
                    help="panels of this run (0: to the end); with --state the run resumes and saves there")
    ap.add_argument("--state", default=None, help="training checkpoint (.npz) to resume from and save into")
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--passes", type=int, default=2, help="plain-cycle error steps per training batch")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=os.path.join(ROOT, "fevc_amd", "smoother_params.json"))
    args = ap.parse_args()

    panels = training_panels(args.panels, args.seed)
    state = np.zeros(3 * COUNT)
    state[:COUNT] = initial_params(args.seed)
    step = 0
    total = args.epochs * args.panels * args.steps
    history = []
    done = 0  # panel visits already trained
    if args.state and os.path.exists(args.state):
        ck = np.load(args.state, allow_pickle=False)
        state, step, done = ck["state"], int(ck["step"]), int(ck["done"])
        history = json.loads(str(ck["history"]))
        log(f"resumed at visit {done} (step {step})")
    visits = [(epoch, spec) for epoch in range(args.epochs) for spec in panels]
    stop = len(visits) if args.count == 0 else min(len(visits), done + args.count)
    for epoch, spec in visits[done:stop]:
        e = engine(spec)
        try:
            lr = args.lr * (0.1 + 0.9 * (1.0 - step / total))
            losses = e.smoother_train(state, step, args.steps, lr,
                                      seed=(args.seed << 20) ^ (epoch << 10) ^ spec.seed,
                                      passes=args.passes)
            dims = e.dims
        finally:
            e.release()
        step += args.steps
        history.append(dict(epoch=epoch, panel=spec.name, workers=spec.workers,
                            hybrid_vertices=dims["hybrid_vertices"], levels=dims["levels"],
                            first_loss=float(losses[0]), last_loss=float(losses[-1])))
        log(f"epoch {epoch} {spec.name}: n={dims['hybrid_vertices']} levels={dims['levels']} "
            f"loss {losses[0]:.4f} -> {losses[-1]:.4f} (lr {lr:.2e})")
    if stop < len(visits):
        np.savez(args.state, state=state, step=step, done=stop, history=json.dumps(history))
        log(f"saved the training state after visit {stop} of {len(visits)} to {args.state}")
        return
    params = state[:COUNT].copy()
    evaluation = []
    for spec in held_out_panels():
        e = engine(spec)
        try:
            e.configure(32, 16, 8)
            e.smoother(1, params)
            st = e.smoother_status()
            evaluation.append(dict(panel=spec.name, workers=spec.workers,
                                   hybrid_vertices=e.dims["hybrid_vertices"],
                                   plain_iterations=st["plain_iterations"],
                                   learned_iterations=st["learned_iterations"], active=st["active"]))
            log(f"held out {spec.name}: calibration iterations {st['plain_iterations']} plain -> "
                f"{st['learned_iterations']} learned")
        finally:
            e.release()
    out = dict(params=[float(v) for v in params],
               meta=dict(count=COUNT, features=FEATURES, hidden=HIDDEN, omega_max=OMEGA_MAX,
                         panels=args.panels, epochs=args.epochs, steps_per_visit=args.steps, lr=args.lr,
                         passes=args.passes, seed=args.seed, adam_steps=step, history=history,
                         held_out=evaluation))
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
