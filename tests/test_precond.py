"""FIRE force preconditioner: FP64 oracle pins (CPU) and device parity (GPU).

The oracle (oracle/precond.py) is pinned against an independent least-squares
construction of the rigid-motion projector and against the algebraic
properties the preconditioner must have (idempotent projectors, a
self-adjoint positive-definite map, identity at zero gain, rigid fields
preserved).  The device kernel is then compared with the oracle evaluated on
the device's own exact forces.
"""
import numpy as np
import pytest

from oracle import precond as pc
from firecore_amd import native


def _rand_group(rng, n, collinear=False):
    if collinear:
        d = rng.normal(size=3)
        return np.outer(rng.normal(size=n), d) + rng.normal(size=3)
    return rng.normal(scale=1.5, size=(n, 3))


@pytest.mark.parametrize("n", [3, 4, 5, 9])
def test_rigid_projector_matches_lstsq(n):
    rng = np.random.default_rng(n)
    x = _rand_group(rng, n)
    F = rng.normal(size=(n, 3))
    a = pc.rigid_project(x, F)
    b = pc.rigid_project_lstsq(x, F)
    assert np.abs(a - b).max() < 1e-10


def test_rigid_projector_collinear_pair():
    rng = np.random.default_rng(7)
    x = _rand_group(rng, 2, collinear=True)
    F = rng.normal(size=(2, 3))
    a = pc.rigid_project(x, F, reg=1e-6)
    b = pc.rigid_project_lstsq(x, F)
    assert np.abs(a - b).max() < 1e-10


def test_projector_idempotent_and_rigid_fields_kept():
    rng = np.random.default_rng(3)
    x = _rand_group(rng, 6)
    F = rng.normal(size=(6, 3))
    P1 = pc.rigid_project(x, F)
    assert np.abs(pc.rigid_project(x, P1) - P1).max() < 1e-10
    t, w = rng.normal(size=3), rng.normal(size=3)
    R = t + np.cross(w, x - x.mean(0))
    assert np.abs(pc.rigid_project(x, R) - R).max() < 1e-10


def _chain(natoms=10, branch=True):
    neighs = -np.ones((natoms, 4), int)

    def bond(a, b):
        neighs[a][list(neighs[a]).index(-1)] = b
        neighs[b][list(neighs[b]).index(-1)] = a

    for i in range(natoms - 1):
        if branch and i == 5:
            bond(2, i + 1)
        else:
            bond(i, i + 1)
    return neighs


def _matrix(x, neighs, nnode, gs, gr):
    n = len(x)
    A = np.zeros((3 * n, 3 * n))
    for k in range(3 * n):
        e = np.zeros(3 * n)
        e[k] = 1
        A[:, k] = pc.precondition(x, e.reshape(n, 3), neighs, nnode, gs, gr, reg=0.0).reshape(-1)
    return A


def test_preconditioner_symmetric_positive_definite():
    rng = np.random.default_rng(11)
    neighs = _chain(10)
    x = rng.normal(scale=2.0, size=(10, 3))
    A = _matrix(x, neighs, 6, rng.uniform(0, 3, 6), 40.0)
    assert np.abs(A - A.T).max() < 1e-9 * np.abs(A).max()
    ev = np.linalg.eigvalsh(0.5 * (A + A.T))
    assert ev.min() > 1 - 1e-9            # I + sum of PSD projectors
    A0 = _matrix(x, neighs, 6, 0.0, 0.0)
    assert np.abs(A0 - np.eye(30)).max() < 1e-12


def test_components_fold():
    neighs = -np.ones((7, 4), int)
    neighs[0, 0], neighs[1, 0] = 1, 0     # {0,1}; 2..6 isolated
    c = pc.components(7, neighs, maxc=4)
    assert list(c) == [0, 0, 1, 2, 3, 3, 3]


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [1, 2])
def test_precond_device_matches_oracle(kind, nacl_grid):
    """Device preconditioned force vs the FP64 oracle applied to the device's
    exact forces, random gains (kind 2: two molecules per system)."""
    native.set_opt()
    native.init_workload(kind, 12, 777, 16, 120, 0.1, "synth:NaCl", 0.1)
    native.set_precond(1, 2.5, 37.0)
    try:
        Ft = native.eval_precond()
        B = native.Batch()
        worst = 0.0
        for s in range(12):
            N, M, a0, n0, v0 = (int(v) for v in B.info[s, :5])
            x = B.atoms[v0:v0 + N, :3].astype(np.float64)
            F = B.aforces[v0:v0 + N, :3].astype(np.float64)
            ref = pc.precondition(x, F, np.asarray(B.neighs[a0:a0 + N]), M, 2.5, 37.0)
            got = Ft[v0:v0 + N, :3]
            worst = max(worst, np.abs(got - ref).max() / (np.abs(ref).max() + 1e-3))
            # pi-vectors keep the exact force
            assert np.array_equal(Ft[v0 + N:v0 + N + M], B.aforces[v0 + N:v0 + N + M])
        assert worst < 2e-5, worst
    finally:
        native.set_precond(0)
