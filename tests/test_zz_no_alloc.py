"""relax_noalloc (the allocation-free variant of relax on the persistent
kernel) against relax itself, through the C-ABI.

Both run the same kernels on the same batch; only the scratch (queue
counters and order, activation slots of the in-kernel network, GridFF window
cache) comes from the library's static device array instead of library
allocations.  Per-system results are deterministic (one workgroup relaxes a
system start to end, no atomics in its arithmetic), so positions, forces,
energies, step counts and statuses must agree bit for bit.  Two calls in a
row, the second on a smaller batch, so that the second call runs over scratch
the first one left behind.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def P(*a):
    print(*a, flush=True)


def _state(native):
    B = native.Batch()
    return (np.array(B.atoms, copy=True), np.array(B.aforces, copy=True), np.array(B.E, copy=True),
            np.array(B.nsteps, copy=True), np.array(B.status, copy=True))


def _compare(native, nsys, seed, nmax, learned):
    native.init_workload(1, nsys, seed, 16, nmax, 0.1, "synth:NaCl", 0.1)
    if learned:
        theta = np.load(os.path.join(HERE, "firecore_amd", "proposal_weights.npy"))
        meta = json.load(open(os.path.join(HERE, "firecore_amd", "proposal_weights.json")))
        native.set_proposal_weights(theta.astype(np.float32))
        native.set_proposal_every(int(meta["every"]))
        native.set_precond(2, meta["gstar"], meta["grigid"])
    else:
        native.set_precond(0)
    n_ref = native.relax(10000, 1e-3, 0, True, True)
    ref = _state(native)
    n_na = native.relax_noalloc(10000, 1e-3, True, True)
    got = _state(native)
    assert n_na == n_ref, (n_na, n_ref)
    for name, a, b in zip(("atoms", "aforces", "E", "nsteps", "status"), ref, got):
        assert np.array_equal(a, b), (name, np.abs(np.asarray(a, float) - np.asarray(b, float)).max())
    assert (ref[4] == native.ST_CONVERGED).all(), np.bincount(ref[4], minlength=4)
    return ref


@pytest.mark.parametrize("learned", [False, True])
def test_noalloc_matches_relax_two_calls(learned):
    from firecore_amd import native
    native.lib()
    native.set_opt(0.1, 0.02, 0.2, 1.1, 0.5, 0.8, 5)
    native.set_force_limit(10.0)
    try:
        P("case: relax_noalloc vs relax, 48 GridFF systems N 16..256, learned=%s" % learned)
        _compare(native, 48, 10001, 256, learned)
        P("case: relax_noalloc vs relax, 12 GridFF systems N 16..120 (smaller, reused scratch)")
        _compare(native, 12, 10101, 120, learned)
    finally:
        native.set_precond(0)
        native.set_force_limit(0.0)
        native.set_opt()


def test_noalloc_rejects_oversize_batch():
    """A batch whose window cache exceeds the static scratch is refused with
    an error, not relaxed over memory that is not there."""
    from firecore_amd import native
    native.lib()
    # 2048 systems of 128..256 atoms: ~390k atoms * 256 B of window cache > 64 MiB
    native.init_workload(1, 2048, 10001, 128, 256, 0.1, "synth:NaCl", 0.1)
    with pytest.raises(RuntimeError):
        native.relax_noalloc(10, 1e-3, True, False)
