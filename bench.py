"""Benchmark: relaxations/sec to Fmax < 1e-3 eV/A on 4096 mixed systems, learned vs plain.

Workload (the headline configuration): 4096 independent synthetic molecules,
N ~ U[16,256] atoms (seeds 10001..14096), each placed 3.0 +- 0.3 A above a
rigid NaCl(001) slab represented by a 56x56x200 cubic B-spline GridFF
(Pauli/London/Coulomb, 0.1 A spacing).  FP32 forces, per-system FIRE
(pyBall/MMFF_multi.py set_opt defaults, per-vec force clamp 10 eV/A as the
reference GPU update kernel), each system stops at its own max|f| < 1e-3
eV/A (exact force), at most 10000 steps.  Synthetic data.

Forms, each timed over the same number of passes:
  plain    -- FIRE on the exact force (the reference dynamics)
  learned  -- FIRE on the preconditioned force whose gains the proposal network
              (firecore_amd/proposal_weights.npy, trained on seeds 1,000,001..)
              re-evaluates inside the relaxation kernel every P steps (P stored
              with the weights, picked on the training systems)
  constant -- the same preconditioner with constant gains (ablation)
One "step" of this benchmark is one full pass of the hot path over the batch:
every system relaxed from its start geometry to convergence (or the step
limit), inputs already resident in HBM.  `value` is the learned form.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NSYS = 4096
NMIN, NMAX = 16, 256
SEED0 = 10001
FCONV = 1e-3
NSTEP_MAX = 10000
FIRE = (0.1, 0.02, 0.2, 1.1, 0.5, 0.8, 5)  # dt_max, dt_min, damp_max, finc, fdec, falpha, minLastNeg
FLIMIT = 10.0
GSTAR_CONST, GRIGID_CONST = 2.0, 63.0
WEIGHTS = os.path.join(ROOT, "firecore_amd", "proposal_weights.npy")
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
GRID_BYTES_ATOM = 64 * 16  # one B-spline GridFF sample: 4x4x4 float4 coefficients
# algorithmic work per unit (FLOP): one unique non-excluded LJ+Coulomb pair,
# one 64-tap B-spline GridFF sample, one bonded node, one FIRE-updated vec;
# preconditioner per node (star projection) and per atom (molecule
# projection); network per atom and evaluation (dense layers 16->64, 128->64
# x2, head 64->2: 2*(16*64 + 2*128*64 + 2*64))
FLOP_PAIR, FLOP_GRID, FLOP_NODE, FLOP_VEC = 38, 740, 400, 40
FLOP_PC_NODE, FLOP_PC_ATOM = 200, 60
FLOP_MLP_ATOM = 2 * (16 * 64 + 2 * 128 * 64 + 2 * 64)
PROP_EVERY = 8


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def topo_counts(B, s):
    """(N, M, V, unique non-excluded pairs) of system s."""
    N, M = int(B.info[s, 0]), int(B.info[s, 1])
    a0 = int(B.info[s, 2])
    ng = B.neighs[a0:a0 + N]
    nb = 0
    n13 = 0
    for k in range(M):
        d = int((ng[k] >= 0).sum())
        n13 += d * (d - 1) // 2
        nb += int(((ng[k] >= 0) & (ng[k] > k)).sum())
    return N, M, N + M, N * (N - 1) // 2 - nb - n13


def cpu_baseline(native, grid, nsample, nthreads):
    import oracle
    idx = np.linspace(0, NSYS - 1, nsample).astype(int)
    systems = [native.gen_workload_system(1, int(i), SEED0, NMIN, NMAX, 0.1, grid) for i in idx]
    fire = tuple(FIRE) + (FLIMIT,)
    t = time.perf_counter()
    steps, E, Fm, _ = oracle.relax_batch(systems, fire, FCONV, NSTEP_MAX, 0, grid=grid, nthreads=nthreads)
    dt = time.perf_counter() - t
    conv = int(np.sum(np.isfinite(Fm) & (Fm < FCONV)))
    return dict(value=conv / dt, unit="relaxations/s", cores=nthreads, kind="port",
                sample="%d of the 4096 systems (evenly spaced indices), FP64 oracle of the reference CPU path, "
                       "plain FIRE, same limits; %d converged in %.1f s" % (nsample, conv, dt))


def measure(native, sync, steps, warmup, tag):
    for w in range(warmup):
        native.relax(NSTEP_MAX, FCONV, 0, True, False)
        log("%s warmup %d: %.1f ms" % (tag, w, native.last_ms()))
    sync()
    t = time.perf_counter()
    dev_ms, class_ms = [], []
    for k in range(steps):
        native.relax(NSTEP_MAX, FCONV, 0, True, False)
        dev_ms.append(native.last_ms())
        class_ms.append([c["ms"] for c in native.class_info()])
        log("%s step %d: %.1f ms" % (tag, k, dev_ms[-1]))
    sync()
    wall = time.perf_counter() - t
    native.lib().download(True, True)
    B = native.Batch()
    status = np.array(B.status)
    nsteps = np.array(B.nsteps)
    conv = int((status == native.ST_CONVERGED).sum())
    ms_per_step = wall * 1e3 / steps
    return dict(ms_per_step=ms_per_step, dev_ms=float(np.mean(dev_ms)), class_ms=class_ms, classes=native.class_info(),
                status=status, nsteps=nsteps, conv=conv, value=conv / (ms_per_step / 1e3), B=B)


def roofline(r, learned, every=PROP_EVERY):
    """Roofline of the dominant kernel (the size class with the longest launch).

    The persistent kernel keeps every system's state in LDS/registers; its
    algorithmic memory traffic is the B-spline GridFF gathers, 64 float4
    coefficients (1024 B) per atom and step from a 10 MB grid that does not
    stay in the per-XCD L2 (a per-atom window cache serves most of them, so
    the measured traffic is lower).  achieved = those algorithmic gather bytes
    of the launch over its measured average duration (HIP events on the
    launch's stream), against the HBM peak; the FP32 work of the same launch
    is reported beside it.
    traffic = L2-to-fabric bytes per launch measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated; profiles/), when recorded."""
    B, steps, classes = r["B"], r["nsteps"], r["classes"]
    flops, gbytes = [], []
    for c in classes:
        f = 0.0
        g = 0.0
        for s in c["order"]:
            N, M, V, P = topo_counts(B, int(s))
            st = float(steps[s] + 1)
            f += st * float(FLOP_PAIR * P + FLOP_GRID * N + FLOP_NODE * M + FLOP_VEC * V)
            g += st * float(GRID_BYTES_ATOM * N)
            if learned:
                f += st * float(FLOP_PC_NODE * M + FLOP_PC_ATOM * N)
                f += (st // every + 1) * float(FLOP_MLP_ATOM * N)
        flops.append(f)
        gbytes.append(g)
    dom = int(np.argmax([c["ms"] for c in classes]))
    dom_ms = float(np.mean([cm[dom] for cm in r["class_ms"]]))
    achieved = gbytes[dom] / (dom_ms * 1e-3) / 1e9
    tflops = flops[dom] / (dom_ms * 1e-3) / 1e12
    kind = 2 if learned else 0
    name = "relax_persistent_kernel<%d, false, %d>" % (classes[dom]["block"], kind)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "relax_persistent_traffic.json")
    if os.path.exists(tf):
        rec = json.load(open(tf)).get(name)
        if rec:
            traffic = rec["bytes_per_launch"]
    return dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                frac=round(achieved / PEAK_HBM_GBS, 4), traffic=traffic, kernel=name,
                kernel_ms=round(dom_ms, 3), kernel_systems=classes[dom]["nsys"],
                algorithmic_bytes=gbytes[dom], fp32_tflops=round(tflops, 3),
                fp32_frac=round(tflops / PEAK_FP32_TFLOPS, 4))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=160)
    ap.add_argument("--forms", default="plain,learned,constant")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="copy a 256 MiB buffer once first (the known-size kernel that scales memory counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the benchmark runs on one GPU; multi-GPU runs (--gpus %d) are not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from firecore_amd import native
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only
        sync = lambda: None  # noqa: E731

    if args.calibrate_copy:
        import torch
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    native.set_opt(*FIRE)
    native.set_force_limit(FLIMIT)
    native.set_conv_mode(0)
    t0 = time.perf_counter()
    native.init_workload(1, NSYS, SEED0, NMIN, NMAX, 0.1, "synth:NaCl", 0.1)
    log("init %.1f s" % (time.perf_counter() - t0))
    grid = native.grid_info()
    forms = args.forms.split(",")
    res = {}
    if "plain" in forms:
        native.set_precond(0)
        res["plain"] = measure(native, sync, args.steps, args.warmup, "plain")
    meta = None
    if "learned" in forms and os.path.exists(WEIGHTS):
        meta = json.load(open(os.path.splitext(WEIGHTS)[0] + ".json"))
        native.set_proposal_weights(np.load(WEIGHTS))
        native.set_proposal_every(int(meta.get("every", PROP_EVERY)))
        native.set_precond(2, meta["gstar"], meta["grigid"])
        res["learned"] = measure(native, sync, args.steps, args.warmup, "learned")
    if "constant" in forms:
        native.set_precond(1, GSTAR_CONST, GRIGID_CONST)
        res["constant"] = measure(native, sync, args.steps, args.warmup, "constant")
    native.set_precond(0)

    head = "learned" if "learned" in res else ("plain" if "plain" in res else list(res)[0])
    r = res[head]
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(native, grid, args.cpu_sample, min(os.cpu_count() or 1, 16))
        log("cpu baseline: %s" % cpu)

    config = {"workload": "4096 uneven molecules N~U[16,256] on NaCl(001) GridFF 56x56x200, FIRE to "
                          "max|f|<1e-3 eV/A, <=10000 steps",
              "form": head, "systems": NSYS, "converged": r["conv"],
              "max_step": int((r["status"] == native.ST_MAXSTEP).sum()),
              "nonfinite": int((r["status"] == native.ST_NAN).sum()),
              "mean_steps": round(float(r["nsteps"].mean()), 1),
              "device_ms_per_step": round(r["dev_ms"], 3)}
    for k, v in res.items():
        config["%s_relax_per_s" % k] = round(v["value"], 2)
        config["%s_mean_steps" % k] = round(float(v["nsteps"].mean()), 1)
        config["%s_converged" % k] = v["conv"]
    if "plain" in res and head != "plain":
        config["%s_over_plain" % head] = round(r["value"] / res["plain"]["value"], 3)
    if meta:
        config["proposal"] = "every %d steps; gain ceilings star %g, molecule %g; trained on %d pairs from seeds %s" % (
            int(meta.get("every", PROP_EVERY)), meta["gstar"], meta["grigid"], meta["pairs"], meta["seeds"])
    out = {
        "metric": "relaxations/sec to Fmax<1e-3 eV/Å, 4096 mixed systems; learned vs plain",
        "value": round(r["value"], 2),
        "unit": "relaxations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": config,
        "roofline": roofline(r, head == "learned", int(meta.get("every", PROP_EVERY)) if meta else PROP_EVERY),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
