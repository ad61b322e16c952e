"""Screening caller scan() (MolWorld_sp3_multi::scan_relaxed,
MolWorld_sp3_multi.h:1242-1296) through the reference-shaped Python API.

Every configuration is the template molecule rotated (rot^T, Mat3::dot_to_T)
and moved to poss[i], atom 0 restrained there (stiffness 1), relaxed on the
device.  Checked against the same systems built by hand and relaxed with
relax(), against the FP64 oracle at the relaxed geometries (force-field
forces, restrained total force below Fconv, energies), and for the batch that
was loaded before being put back unchanged.
"""
import numpy as np
import pytest

from conftest import FIRE_MULTI

pytestmark = pytest.mark.gpu


def _rot(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def test_scan_relaxed_configurations(orc):
    from firecore_amd import MMFF_multi as mm
    from firecore_amd import native as nat
    print("case: scan", flush=True)
    rng = np.random.default_rng(21)
    mm.init(nSys_=6, xyz_name="synth:natoms=24,seed=0,sigma=0.1", nPBC=(0, 0, 0))
    mm.getBuffs()
    nat.set_opt(*FIRE_MULTI)
    before = nat.Batch().positions(3).copy()
    nconf = 10
    poss = rng.uniform(-2, 2, (nconf, 3))
    rots = np.array([_rot(rng) for _ in range(nconf)])
    Es, F, X = mm.scan(poss, rots, bF=True, bP=True, bRelax=True, niter_max=8000, Fconv=1e-3, Flim=10.0)
    assert np.isfinite(Es).all() and F.shape == (nconf, 24, 3) and X.shape == (nconf, 24, 3)
    # the batch that was loaded before is back
    B = nat.Batch()
    assert B.nsys == 6 and np.array_equal(B.positions(3), before)
    # the same configurations by hand: template moved, atom 0 restrained, relax()
    tmpl = nat.gen_molecule(24, 0)
    tm = np.array(tmpl["apos"])
    mols = []
    for i in range(nconf):
        m = dict(tmpl)
        a = tm @ rots[i]           # row vectors: rot^T v
        a[:24] += poss[i]
        m["apos"] = a
        mols.append(m)
    nat.init_systems(mols)
    B = nat.Batch()
    for i in range(nconf):
        a0 = int(B.info[i, 2])
        B.constr[a0] = list(poss[i]) + [1.0]
        B.constrK[a0] = [1.0, 1.0, 1.0, 1.0]
    nat.upload(False, False, False, True)
    nat.set_force_limit(10.0)
    try:
        nat.relax(8000, 1e-3, 0, True, True)
    finally:
        nat.set_force_limit(0.0)
    B = nat.Batch()
    assert (B.status == nat.ST_CONVERGED).all()
    for i in range(nconf):
        pos = B.positions(i)
        assert np.array_equal(pos[:24].astype(np.float32), X[i].astype(np.float32)), i
        assert abs(B.E[i] - Es[i]) == 0.0, i
        m = dict(tmpl, apos=pos)
        fo, _, _, Eo = orc.eval(m)
        assert np.abs(fo[:24] - F[i]).max() < 1e-4, i          # force-field force, restraint removed
        assert abs(Eo - Es[i]) < 1e-4, i                         # force-field energy
        c = -np.ones((24, 4))
        c[0] = list(poss[i]) + [1.0]
        k = -np.ones((24, 4))
        k[0] = 1.0
        orc.set_restraints(24, c, k)
        try:
            ft = orc.eval(m)[0]
        finally:
            orc.set_restraints(0)
        assert np.sqrt((ft ** 2).sum(1)).max() < 1.01e-3, i      # relaxed under the restraint
        assert np.linalg.norm(pos[0] - poss[i]) < 0.5, i
