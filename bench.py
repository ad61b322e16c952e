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

Before the timed passes each form's launch plan (queue-order cost model,
scheduling mode, workgroup widths) is tuned by firecore_amd/autotune.py on
4096 systems of seeds 90001.. (disjoint from the benchmark batch); the timed
passes never use step counts of earlier passes over the benchmark batch.
A second leg relaxes the screening workload: 65,536 molecule pairs (8-64
atoms each, seeds 20001..), learned form, at most 3000 steps, with its launch
plan tuned on a batch of the same size from seeds 190001...

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu] [--no-screening]
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
CALIB_SEED0 = 90001  # calibration batch of the launch plans (disjoint from SEED0..)
FCONV = 1e-3
NSTEP_MAX = 10000
FIRE = (0.1, 0.02, 0.2, 1.1, 0.5, 0.8, 5)  # dt_max, dt_min, damp_max, finc, fdec, falpha, minLastNeg
FLIMIT = 10.0
# best constant gains of a sweep over star {1,2,3} x molecule {31,63,127} on
# 2048 training-seed systems (1000001..): shortest pass, every system converged
GSTAR_CONST, GRIGID_CONST = 2.0, 127.0
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
SCREEN_NSYS, SCREEN_SEED0, SCREEN_STEPS = 65536, 20001, 3000
SCREEN_CALIB_SEED0 = 190001  # calibration batch of the screening plan (disjoint from 20001.. and 10001..)


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


def host_info():
    """CPU model and counts of the host the baseline runs on."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return dict(model=model, nproc=os.cpu_count() or 1, affinity=affinity)


def cpu_baseline(native, grid, nsample, nthreads, reps=5, nsys=NSYS, reps32=3):
    """The oracle of the reference CPU path (plain FIRE, one system per
    thread, dynamic schedule) on an evenly spaced sample of the workload:
    FP64 (the reported value), one warm-up run, then the median of `reps`
    timed runs; and the same restatement built in single precision (every
    array and operation FP32), median of `reps32` after a warm-up."""
    import oracle
    idx = np.linspace(0, nsys - 1, nsample).astype(int)
    systems = [native.gen_workload_system(1, int(i), SEED0, NMIN, NMAX, 0.1, grid) for i in idx]
    fire = tuple(FIRE) + (FLIMIT,)

    def timed(precision, nrep):
        rates, convs = [], []
        for r in range(nrep + 1):
            t = time.perf_counter()
            steps, E, Fm, _ = oracle.relax_batch(systems, fire, FCONV, NSTEP_MAX, 0, grid=grid, nthreads=nthreads,
                                                 precision=precision)
            dt = time.perf_counter() - t
            conv = int(np.sum(np.isfinite(Fm) & (Fm < FCONV)))
            if r > 0:
                rates.append(conv / dt)
                convs.append(conv)
            log("cpu baseline %s run %d: %d converged in %.2f s" % (precision, r, conv, dt))
        return rates, convs

    rates, convs = timed("f64", reps)
    rates32, convs32 = timed("f32", reps32)
    h = host_info()
    return dict(value=float(np.median(rates)), unit="relaxations/s", cores=nthreads, kind="port",
                sample="%d of the %d systems (evenly spaced indices), FP64 oracle of the reference CPU path, "
                       "plain FIRE, same limits; %d converged; median of %d runs after 1 warm-up (%.2f..%.2f/s)"
                       % (nsample, nsys, convs[0], reps, min(rates), max(rates)),
                fp32=dict(value=float(np.median(rates32)), converged=convs32[0],
                          note="the same restatement in single precision, median of %d runs after 1 warm-up "
                               "(%.2f..%.2f/s); scalar code, so FP32 gains little over FP64"
                               % (reps32, min(rates32), max(rates32))),
                host=h["model"], host_nproc=h["nproc"], host_affinity=h["affinity"],
                threads_note="one system per thread on %d threads: this box's CPU share per GPU" % nthreads)


def measure(native, sync, steps, warmup, tag, nstep_max=NSTEP_MAX, path=0):
    """Time `steps` relaxation passes after `warmup` untimed ones.  The queue
    order comes from the autotuned cost model (no step counts of earlier
    passes), so every pass, the first included, schedules the same way.
    path 0: the persistent kernel; 1: the split kernels (one launch per phase
    and step)."""
    first_ms = None
    for w in range(warmup):
        native.relax(nstep_max, FCONV, path, True, False)
        log("%s warmup %d: %.1f ms" % (tag, w, native.last_ms()))
        first_ms = native.last_ms() if first_ms is None else first_ms
    sync()
    t = time.perf_counter()
    dev_ms, class_ms = [], []
    for k in range(steps):
        native.relax(nstep_max, FCONV, path, True, False)
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
    return dict(ms_per_step=ms_per_step, dev_ms=float(np.mean(dev_ms)), first_ms=first_ms, class_ms=class_ms,
                classes=native.class_info(),
                status=status, nsteps=nsteps, conv=conv, value=conv / (ms_per_step / 1e3), B=B)


def roofline(r, learned, every=PROP_EVERY, plan=None):
    """Roofline of the relaxation kernel over one pass.

    A pass launches relax_persistent_kernel once per size class (the same code
    at a few workgroup widths) on its own stream; the launches overlap and
    share the device, and persistent workgroups leave only when their class
    queue is empty, so a class's systems start when the previous launch
    drains.  The unit is therefore the pass: achieved = the algorithmic FLOPs
    of all systems of the pass (per-unit counts in FLOP_*) over the pass's
    device time, measured with HIP events on the stream that forks and joins
    the class launches.  The kernel keeps every system in LDS and registers
    and is bound by FP32 arithmetic (the half-ring pair loop, the B-spline
    GridFF sample, the bonded terms), so the roof is the chip's FP32 peak
    (157.3 TFLOP/s, a packed-FP32 figure; this build never emits packed FP32,
    so half of it is the reachable ceiling).  traffic = the HBM-side bytes of
    all class launches, measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, copy-calibrated; profiles/), hbm_frac its share of 8 TB/s over
    the pass.  grid_gather_bytes: the algorithmic B-spline coefficient bytes
    (1 KiB per atom and step), most of which the per-atom window cache serves
    without touching the grid.  `launch` gives the same figures for the class
    launch that ends last, over its own HIP-event duration on its stream.
    traffic_plan: the launch plan the counters were measured with (the bytes
    of a class launch depend on which systems the plan gives it)."""
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
    kind = 2 if learned else 0
    names = ["relax_persistent_kernel<%d, false, %d>" % (c["block"], kind) for c in classes]
    rec = {}
    tf = os.path.join(ROOT, "profiles", "relax_persistent_traffic.json")
    if os.path.exists(tf):
        rec = json.load(open(tf))
    pass_ms = r["dev_ms"]
    pass_flop = float(np.sum(flops))
    pass_tflops = pass_flop / (pass_ms * 1e-3) / 1e12
    live = [k for k, c in enumerate(classes) if c["nsys"] > 0]
    traffic = float(sum(rec[names[k]]["bytes_per_launch"] for k in live)) if all(names[k] in rec for k in live) \
        else None
    dom = int(np.argmax([c["ms"] for c in classes]))
    dom_ms = float(np.mean([cm[dom] for cm in r["class_ms"]]))
    dom_traffic = rec[names[dom]]["bytes_per_launch"] if names[dom] in rec else None
    launch = dict(kernel=names[dom], ms=round(dom_ms, 3), systems=classes[dom]["nsys"],
                  algorithmic_flop=flops[dom], achieved=round(flops[dom] / (dom_ms * 1e-3) / 1e12, 3),
                  frac=round(flops[dom] / (dom_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4), traffic=dom_traffic)
    return dict(bound="mfma", achieved=round(pass_tflops, 3), peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                frac=round(pass_tflops / PEAK_FP32_TFLOPS, 4), traffic=traffic,
                pipe="FP32 vector ALU (the step loop; the learned form's in-kernel network adds BF16 MFMA "
                     "work of under 1 % of the kernel's cycles); peak = the FP32 figure",
                kernel="relax_persistent_kernel, the %d size-class launches of one pass (%s)"
                       % (len(live), ", ".join(names[k] for k in live)),
                kernel_ms=round(pass_ms, 3), kernel_systems=int(sum(classes[k]["nsys"] for k in live)),
                algorithmic_flop=pass_flop,
                hbm_frac=round(traffic / (pass_ms * 1e-3) / (PEAK_HBM_GBS * 1e9), 4) if traffic else None,
                grid_gather_bytes=float(np.sum(gbytes)), launch=launch,
                traffic_plan=dict(rec.get("_profiled_plans", {}).get("learned" if learned else "plain", {}),
                                  same_as_this_pass=bool(plan) and all(
                                      rec.get("_profiled_plans", {}).get("learned" if learned else "plain", {}).get(k)
                                      == plan.get(k) for k in ("sched", "class_mask"))))


def load_trained(native):
    meta = json.load(open(os.path.splitext(WEIGHTS)[0] + ".json"))
    native.set_proposal_weights(np.load(WEIGHTS))
    native.set_proposal_every(int(meta.get("every", PROP_EVERY)))
    return meta


def form_setup(native, form, meta):
    if form == "plain":
        return lambda: native.set_precond(0)
    if form == "constant":
        return lambda: native.set_precond(1, GSTAR_CONST, GRIGID_CONST)

    def learned():
        load_trained(native)
        native.set_precond(2, meta["gstar"], meta["grigid"])
    return learned


def screening_leg(native, sync, autotune, meta, nsys, steps):
    """Molecule-pair screening workload: nsys systems of two molecules of 8-64
    atoms (random relative orientation, 4-8 A apart) over the same GridFF,
    learned form, at most 3000 steps."""
    setup = form_setup(native, "learned", meta)
    # calibration batch of the same size from disjoint seeds (queueing, and so
    # the best widths, depend on the batch size)
    plan = autotune.calibrate(setup, kind=2, nsys=nsys, seed0=SCREEN_CALIB_SEED0, nmin=8, nmax=64,
                              nstep_max=SCREEN_STEPS, log=log)
    t0 = time.perf_counter()
    native.init_workload(2, nsys, SCREEN_SEED0, 8, 64, 0.1, "synth:NaCl", 0.1)
    log("screening init %.1f s" % (time.perf_counter() - t0))
    setup()
    autotune.apply(plan)
    r = measure(native, sync, steps, 1, "screening", nstep_max=SCREEN_STEPS)
    return dict(workload="%d molecule pairs (8-64 atoms each) on the NaCl GridFF, learned form, FIRE to "
                         "max|f|<1e-3 eV/A, <=%d steps (seeds %d..)" % (nsys, SCREEN_STEPS, SCREEN_SEED0),
                relax_per_s=round(r["value"], 2), converged=r["conv"], systems=nsys,
                max_step=int((r["status"] == native.ST_MAXSTEP).sum()),
                mean_steps=round(float(r["nsteps"].mean()), 1), ms_per_pass=round(r["ms_per_step"], 3),
                first_pass_ms=round(r["first_ms"], 3) if r["first_ms"] else None, plan=plan)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=48)
    ap.add_argument("--forms", default="plain,learned,constant")
    ap.add_argument("--no-screening", action="store_true")
    ap.add_argument("--screening-systems", type=int, default=SCREEN_NSYS)
    ap.add_argument("--path", choices=("persistent", "split"), default="persistent",
                    help="relaxation kernels: the persistent kernel (default) or the split per-phase kernels "
                         "(a profiling mode: no launch plan, no roofline)")
    ap.add_argument("--systems", type=int, default=NSYS, help="systems of the workload (default %d)" % NSYS)
    ap.add_argument("--plans-out", help="write the calibrated launch plans (JSON) to this file")
    ap.add_argument("--plans-in", help="use the launch plans of this file instead of calibrating (counter runs, "
                                       "which serialise the launches and would calibrate differently)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="copy a 256 MiB buffer once first (the known-size kernel that scales memory counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the benchmark runs on one GPU; multi-GPU runs (--gpus %d) are not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from firecore_amd import autotune, native
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
    forms = [f for f in args.forms.split(",") if f in ("plain", "learned", "constant")]
    meta = json.load(open(os.path.splitext(WEIGHTS)[0] + ".json")) if os.path.exists(WEIGHTS) else None
    if meta is None:
        forms = [f for f in forms if f != "learned"]
    # launch plans, tuned per form on systems of seeds disjoint from the benchmark's
    path = 1 if args.path == "split" else 0
    plans = json.load(open(args.plans_in)) if args.plans_in else {}
    # one calibration batch of the benchmark's size and distribution from
    # disjoint seeds, tuned for every form (queueing in a pass depends on the
    # batch size, so a smaller batch picks other widths)
    calib_init = True
    for f in forms:
        if path == 1:
            plans[f] = None
            continue
        if f in plans:
            continue
        plans[f] = autotune.calibrate(form_setup(native, f, meta), kind=1, nsys=args.systems, seed0=CALIB_SEED0,
                                      nmin=NMIN, nmax=NMAX, nstep_max=NSTEP_MAX, log=log, init=calib_init)
        calib_init = False
    if args.plans_out:
        with open(args.plans_out, "w") as fh:
            json.dump(plans, fh)
    t0 = time.perf_counter()
    native.init_workload(1, args.systems, SEED0, NMIN, NMAX, 0.1, "synth:NaCl", 0.1)
    log("init %.1f s" % (time.perf_counter() - t0))
    grid = native.grid_info()
    res = {}
    for f in forms:
        form_setup(native, f, meta)()
        if plans[f]:
            autotune.apply(plans[f])
        res[f] = measure(native, sync, args.steps, args.warmup, f, path=path)
    native.set_precond(0)

    head = "learned" if "learned" in res else ("plain" if "plain" in res else list(res)[0])
    r = res[head]
    config = {"workload": "%d uneven molecules N~U[16,256] on NaCl(001) GridFF 56x56x200, FIRE to "
                          "max|f|<1e-3 eV/A, <=10000 steps" % args.systems,
              "form": head, "systems": args.systems, "converged": r["conv"],
              "max_step": int((r["status"] == native.ST_MAXSTEP).sum()),
              "nonfinite": int((r["status"] == native.ST_NAN).sum()),
              "mean_steps": round(float(r["nsteps"].mean()), 1),
              "device_ms_per_step": round(r["dev_ms"], 3),
              "first_pass_ms": round(r["first_ms"], 3) if r["first_ms"] else None,
              "scheduling": "queue order from the autotuned cost model steps(N)*N^2, scheduling mode and class widths "
                            "tuned on %d systems of seeds %d..; no step counts of earlier passes of the benchmark "
                            "batch are used" % (args.systems, CALIB_SEED0),
              "plan": plans[head]}
    for k, v in res.items():
        config["%s_relax_per_s" % k] = round(v["value"], 2)
        config["%s_mean_steps" % k] = round(float(v["nsteps"].mean()), 1)
        config["%s_converged" % k] = v["conv"]
    if "plain" in res and head != "plain":
        config["%s_over_plain" % head] = round(r["value"] / res["plain"]["value"], 3)
    if "constant" in res and head == "learned":
        config["learned_over_constant"] = round(r["value"] / res["constant"]["value"], 3)
    if meta:
        config["proposal"] = "every %d steps; gain ceilings star %g, molecule %g; trained on %d pairs from seeds %s" % (
            int(meta.get("every", PROP_EVERY)), meta["gstar"], meta["grigid"], meta["pairs"], meta["seeds"])
    if path == 1:
        config["path"] = "split kernels (profiling mode)"
    rl = roofline(r, head == "learned", int(meta.get("every", PROP_EVERY)) if meta else PROP_EVERY,
                  plans.get(head)) if path == 0 else None
    if not args.no_screening and meta:
        config["screening"] = screening_leg(native, sync, autotune, meta, args.screening_systems, 1)
    native.set_precond(0)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(native, grid, args.cpu_sample, min(host_info()["affinity"], 16), nsys=args.systems)
        log("cpu baseline: %s" % cpu)
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
        "roofline": rl,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
