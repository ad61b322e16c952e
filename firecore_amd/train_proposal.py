"""Train the proposal network on the GPU and write its parameters.

    python -m firecore_amd.train_proposal [--systems 4096] [--iters 4000] [--out PATH]

Training data: plain-FIRE relaxations (the library's persistent kernel, no
preconditioner) of synthetic molecules N ~ U[16,256] placed over the NaCl(001)
GridFF, generated from seeds 1,000,001.. (disjoint from the 10,001.. seeds of
the benchmark workload).  Positions are recorded every `--every` steps; every
training step draws one recorded geometry per system and minimises the mean
per-atom energy change of one preconditioned steepest-descent step
(E(x0 + eta*F~) - E(x0)) / N with respect to the network parameters.  The
forward pass, the force evaluations, the backward pass and the Adam update all
run as the library's device kernels; this script only drives them.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from firecore_amd import native  # noqa: E402

WEIGHTS = os.path.join(HERE, "proposal_weights.npy")
FIRE = (0.1, 0.02, 0.2, 1.1, 0.5, 0.8, 5)
GSTAR, GRIGID = 3.0, 127.0        # default gain ceilings of the network's outputs
G0_STAR, G0_RIGID = 2.0, 63.0     # initial gains (the constant-gain setting)
TRAIN_SEED0 = 1000001


def initial_params(seed=0, gstar=GSTAR, grigid=GRIGID, g0star=G0_STAR, g0rigid=G0_RIGID):
    """He-initialised dense layers, a small head whose biases start the gains
    at (g0star, g0rigid) under the ceilings (gstar, grigid)."""
    rng = np.random.default_rng(seed)
    nh, nin = 64, 16
    W1 = rng.normal(scale=np.sqrt(2.0 / nin), size=(nh, nin))
    W2 = rng.normal(scale=np.sqrt(2.0 / (2 * nh)), size=(nh, 2 * nh))
    W3 = rng.normal(scale=np.sqrt(2.0 / (2 * nh)), size=(nh, 2 * nh))
    W4 = rng.normal(scale=0.01, size=(2, nh))
    logit = lambda p: np.log(p / (1.0 - p))  # noqa: E731
    b4 = np.array([logit(g0star / gstar), logit(g0rigid / grigid)])
    z = np.zeros(nh)
    return np.concatenate([W1.ravel(), z, W2.ravel(), z, W3.ravel(), z, W4.ravel(), b4]).astype(np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=4096)
    ap.add_argument("--nmin", type=int, default=16)
    ap.add_argument("--nmax", type=int, default=256)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--nsnap", type=int, default=100)
    ap.add_argument("--iters", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eta", type=float, default=0.02)
    ap.add_argument("--dmax", type=float, default=0.5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--gstar", type=float, default=GSTAR)
    ap.add_argument("--grigid", type=float, default=GRIGID)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--every-candidates", type=int, nargs="+", default=[32, 64, 128, 256])
    args = ap.parse_args()

    t0 = time.time()
    native.set_opt(*FIRE)
    native.set_force_limit(10.0)
    native.set_conv_mode(0)
    native.init_workload(1, args.systems, TRAIN_SEED0, args.nmin, args.nmax, 0.1, "synth:NaCl", 0.1)
    print("init %.1f s" % (time.time() - t0), flush=True)
    native.set_precond(0)
    native.record_snapshots(args.every, args.nsnap)
    native.relax(args.every * args.nsnap, 1e-3, 0, True, False)
    native.lib().download(True, True)
    B = native.Batch()
    steps = np.array(B.nsteps)
    pairs = int(np.minimum(args.nsnap, steps // args.every + 1).sum())
    print("trajectories: %d systems, mean %.0f steps, %d snapshot geometries (%.1f s)"
          % (args.systems, steps.mean(), pairs, time.time() - t0), flush=True)

    native.set_proposal_weights(initial_params(args.seed, args.gstar, args.grigid))
    native.set_precond(2, args.gstar, args.grigid)
    t1 = time.time()
    hist = []
    chunk = 50
    for k in range(0, args.iters, chunk):
        n = min(chunk, args.iters - k)
        h = native.train_proposal(n, args.lr, args.eta, args.dmax, seed=args.seed * 1000003 + k)
        hist.extend(h.tolist())
        print("iter %4d  loss %.6f eV/atom  (%.1f s)" % (k + n, float(np.mean(h[-10:])), time.time() - t1),
              flush=True)
    theta = native.get_proposal_weights()
    native.record_snapshots(1, 0)
    # proposal period: the fastest of the candidates on the training systems,
    # measured with the trained weights (second pass, longest-first order)
    native.set_proposal_weights(theta)
    native.set_precond(2, args.gstar, args.grigid)
    best = None
    for every in args.every_candidates:
        native.set_proposal_every(every)
        native.relax(10000, 1e-3, 0, True, False)  # step counts for the longest-first queue order
        native.relax(10000, 1e-3, 0, True, False)
        ms = native.last_ms()
        native.lib().download(True, True)
        conv = int((np.array(native.Batch().status) == native.ST_CONVERGED).sum())
        rate = conv / ms * 1e3
        print("period %d: %.1f ms, %d converged, %.0f relaxations/s (training systems)" % (every, ms, conv, rate),
              flush=True)
        if best is None or rate > best[1]:
            best = (every, rate)
    native.set_precond(0)
    np.save(args.out, theta.astype(np.float32))
    meta = dict(systems=args.systems, seeds="%d..%d" % (TRAIN_SEED0, TRAIN_SEED0 + args.systems - 1),
                pairs=pairs, iters=args.iters, lr=args.lr, eta=args.eta, dmax=args.dmax, gstar=args.gstar,
                grigid=args.grigid, loss_first=float(np.mean(hist[:10])), loss_last=float(np.mean(hist[-10:])),
                train_seconds=round(time.time() - t1, 2), every=best[0],
                every_note="proposal period picked on the training systems from %s" % args.every_candidates)
    with open(os.path.splitext(args.out)[0] + ".json", "w") as fh:
        json.dump(meta, fh, indent=1)
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
