"""Positional restraints and bounding walls of the update kernel
(updateAtomsMMFFf4, relax_multi.cl:997-1009): per atom constr = (x, y, z, on)
with per-axis stiffness constrK, per system walls (lower corner, upper
corner, per-axis stiffness).  They are part of the exact force.

CPU: the FP64 oracle's restraint and wall forces are the negative gradient of
E_r = sum 1/2 k (c - p)^2 + sum 1/2 k_w (bound - p)^2 (finite differences).
GPU: device forces with restraints and walls against the oracle; a restrained
relaxation converges on the total force and keeps the anchored atom near its
anchor.
"""
import numpy as np
import pytest

from conftest import FIRE_MULTI


def _restraints(m, rng):
    N = m["natoms"]
    constr = -np.ones((N, 4))
    constrK = -np.ones((N, 4))
    pick = rng.choice(N, 3, replace=False)
    for i in pick:
        constr[i, :3] = m["apos"][i] + rng.normal(scale=0.3, size=3)
        constr[i, 3] = 1.0
        constrK[i, :3] = rng.uniform(0.5, 2.0, 3)
    constrK[pick[0], 1] = -1.0  # a negative stiffness is no stiffness on that axis
    z = m["apos"][:N, 2]
    bbox = np.array([-1e8, -1e8, z.min() + 0.4, 1e8, 1e8, z.max() - 0.4, 0.0, 0.0, 3.0])
    return constr, constrK, bbox, pick


def test_oracle_restraint_forces_are_gradients(orc):
    from firecore_amd import native
    rng = np.random.default_rng(5)
    m = native.gen_molecule(30, 7, 0.1, 8)
    N = m["natoms"]
    constr, constrK, bbox, pick = _restraints(m, rng)
    f0 = orc.eval(m)[0]
    orc.set_restraints(N, constr, constrK, bbox)
    try:
        f1 = orc.eval(m)[0]
    finally:
        orc.set_restraints(0)
    df = (f1 - f0)[:N]

    def Er(pos):
        e = 0.0
        for i in range(N):
            if constr[i, 3] > 0:
                e += 0.5 * (np.maximum(constrK[i, :3], 0) * (constr[i, :3] - pos[i]) ** 2).sum()
            for d in range(3):
                if bbox[6 + d] > 0:
                    if pos[i, d] < bbox[d]:
                        e += 0.5 * bbox[6 + d] * (bbox[d] - pos[i, d]) ** 2
                    elif pos[i, d] > bbox[3 + d]:
                        e += 0.5 * bbox[6 + d] * (bbox[3 + d] - pos[i, d]) ** 2
        return e
    pos = np.array(m["apos"][:N])
    h = 1e-6
    fd = np.zeros((N, 3))
    for i in range(N):
        for d in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, d] += h
            p2[i, d] -= h
            fd[i, d] = -(Er(p1) - Er(p2)) / (2 * h)
    assert np.abs(df - fd).max() < 1e-6, np.abs(df - fd).max()
    assert np.abs(df).max() > 0.1  # the restraints and walls act
    assert df[pick[0], 1] == 0.0


def _with_restraints(nat, mols, rng):
    nat.init_systems(mols)
    B = nat.Batch()
    out = []
    for s, m in enumerate(mols):
        N, a0 = m["natoms"], int(B.info[s, 2])
        constr, constrK, bbox, pick = _restraints(m, rng)
        B.constr[a0:a0 + N] = constr
        B.constrK[a0:a0 + N] = constrK
        B.bboxes[s] = np.array([list(bbox[0:3]) + [0], list(bbox[3:6]) + [0], list(bbox[6:9]) + [0]])
        out.append((constr, constrK, bbox, pick))
    nat.upload(False, False, False, True)
    return out


@pytest.mark.gpu
def test_restraint_forces_on_device(orc):
    from firecore_amd import native as nat
    print("case: restraint forces", flush=True)
    rng = np.random.default_rng(11)
    mols = [nat.gen_molecule(n, s, 0.1, s + 9) for n, s in [(24, 51), (70, 52), (180, 53)]]
    rs = _with_restraints(nat, mols, rng)
    for path in (1, 0):  # split evaluation, then a zero-step persistent pass
        if path == 1:
            nat.eval_forces()
        else:
            nat.relax(0, 1e-12, 0, True, True)
        B = nat.Batch()
        for s, (m, (constr, constrK, bbox, _)) in enumerate(zip(mols, rs)):
            orc.set_restraints(m["natoms"], constr, constrK, bbox)
            try:
                fo, _, _, Eo = orc.eval(m)
            finally:
                orc.set_restraints(0)
            d = np.abs(B.forces(s) - fo).max()
            assert d < 1e-4 + 5e-6 * np.abs(fo).max(), (path, s, d)


@pytest.mark.gpu
def test_restrained_relaxation(orc):
    from firecore_amd import native as nat
    print("case: restrained relaxation", flush=True)
    rng = np.random.default_rng(12)
    nat.set_opt(*FIRE_MULTI)
    nat.set_force_limit(10.0)
    try:
        mols = [nat.gen_molecule(n, s, 0.1, s + 9) for n, s in [(24, 61), (90, 62), (200, 63)]]
        rs = _with_restraints(nat, mols, rng)
        for path in (0, 1):
            nat.relax(20000, 1e-3, path, True, True)
            B = nat.Batch()
            assert (B.status == nat.ST_CONVERGED).all(), (path, B.status)
            for s, (m, (constr, constrK, bbox, pick)) in enumerate(zip(mols, rs)):
                N = m["natoms"]
                orc.set_restraints(N, constr, constrK, bbox)
                try:
                    fo = orc.eval(m, B.positions(s))[0]
                finally:
                    orc.set_restraints(0)
                assert np.sqrt((fo ** 2).sum(1)).max() < 1.01e-3, (path, s)
                pos = B.positions(s)[:N]
                # the walls hold: nobody far outside the z band
                assert pos[:, 2].min() > bbox[2] - 0.5 and pos[:, 2].max() < bbox[5] + 0.5
    finally:
        nat.set_force_limit(0.0)
