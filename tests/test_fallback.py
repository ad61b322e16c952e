"""The learned form against constant gains, and its FIRE fallback (neither has
a counterpart in the reference, which has no learned component).

Fallback: a system whose max|F| at `K` proposal points in a row is no lower
than its best one so far continues with plain FIRE (zero gains, no further
network evaluations); set_proposal_stall(K), 0 = never (the default).
GPU: with the fallback on, every system still reaches an FP64-oracle minimum
(max|F| < 1.01e-3 eV/A over all vecs at the device geometry, |dE| < 1e-4 eV)
in both learned modes, and the fallen-back systems are flagged; with it off
none is.  The trained network needs fewer FIRE steps than the best
constant-gain setting (star 2, molecule 127, picked by a sweep on the
training seeds) on held-out seeds.
"""
import json
import os

import numpy as np
import pytest

from firecore_amd import native

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIRE_MULTI = (0.1, 0.02, 0.2, 1.1, 0.5, 0.8, 5)
BEST_CONST = (2.0, 127.0)


def P(*a):
    print(*a, flush=True)


def _trained():
    th = np.load(os.path.join(ROOT, "firecore_amd", "proposal_weights.npy"))
    meta = json.load(open(os.path.join(ROOT, "firecore_amd", "proposal_weights.json")))
    return th, meta


def _learned(meta, th):
    native.set_proposal_weights(th)
    native.set_proposal_every(int(meta["every"]))
    native.set_precond(2, meta["gstar"], meta["grigid"])


@pytest.mark.parametrize("inkernel", [1, 0])
def test_fallback_systems_reach_oracle_minima(orc, inkernel):
    from test_proposal import _check_relaxed_against_oracle
    th, meta = _trained()
    native.set_opt(*FIRE_MULTI)
    native.set_force_limit(10.0)
    seed0, n, nmin, nmax = 30001, 40, 16, 256
    native.init_workload(1, n, seed0, nmin, nmax, 0.1, "synth:NaCl", 0.1)
    _learned(meta, th)
    old = native.set_learned_inkernel(inkernel)
    try:
        native.set_proposal_stall(0)
        native.relax(20000, 1e-3, 0, True, True)
        assert not native.proposal_fallback().any()
        P("case: learned form with the FIRE fallback after 2 windows, inkernel", inkernel)
        native.set_proposal_stall(2)
        native.relax(20000, 1e-3, 0, True, True)
        fb = native.proposal_fallback()
        B = native.Batch()
        assert (np.array(B.status) == native.ST_CONVERGED).all(), np.array(B.status)
        assert fb.sum() >= 5, fb.sum()  # most systems plateau for two windows at some point
        _check_relaxed_against_oracle(orc, B, 1, seed0, nmin, nmax, range(n))
        # the flags describe the last relaxation only: a constant-gain one clears them
        native.set_precond(1, *BEST_CONST)
        native.relax(20000, 1e-3, 0, True, True)
        assert not native.proposal_fallback().any()
    finally:
        native.set_proposal_stall(0)
        native.set_learned_inkernel(old)
        native.set_precond(0)
        native.set_force_limit(0.0)
        native.set_opt()


def test_trained_network_beats_best_constant_gains():
    """Held-out seeds (40001.., disjoint from training 1000001.. and from the
    benchmark's 10001..): fewer mean FIRE steps than the best constant gains."""
    th, meta = _trained()
    native.set_opt(*FIRE_MULTI)
    native.set_force_limit(10.0)
    native.init_workload(1, 1024, 40001, 16, 256, 0.1, "synth:NaCl", 0.1)
    try:
        steps = {}
        for form in ("constant", "learned"):
            P("case: held-out relaxation,", form)
            if form == "constant":
                native.set_precond(1, *BEST_CONST)
            else:
                _learned(meta, th)
            native.relax(10000, 1e-3, 0, True, True)
            B = native.Batch()
            st = np.array(B.status)
            steps[form] = np.array(B.nsteps, float)
            assert (st == native.ST_CONVERGED).mean() > 0.995, (form, (st != native.ST_CONVERGED).sum())
        ratio = steps["learned"].mean() / steps["constant"].mean()
        P("mean steps learned %.1f constant %.1f ratio %.3f" % (steps["learned"].mean(), steps["constant"].mean(),
                                                                  ratio))
        assert ratio < 0.98, ratio
    finally:
        native.set_precond(0)
        native.set_force_limit(0.0)
        native.set_opt()
