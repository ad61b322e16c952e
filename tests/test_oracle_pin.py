"""Pins the FP64 oracle (oracle/oracle.cpp) against the reference.

* against the reference's own CPU classes built here (oracle/_ref), on fresh
  random molecules -- skipped where /root/reference was not available;
* against the golden fixtures in tests/golden/ written from that build.
"""
import os

import numpy as np
import pytest

from firecore_amd import native

GOLD = os.path.join(os.path.dirname(__file__), "golden")
from conftest import FIRE_REF  # noqa: E402


def _mol(z, k):
    keys = ("apos", "REQs", "neighs", "apars", "bLs", "bKs", "Ksp", "Kpp", "elem")
    m = {key: z["m%d_%s" % (k, key)] for key in keys}
    m["natoms"] = int(z["m%d_natoms" % k])
    m["nnode"] = int(z["m%d_nnode" % k])
    return m


def _grid(z):
    return dict(n=tuple(int(x) for x in z["gn"]), pos0=z["gpos0"], step=z["gstep"], alpha=float(z["galpha"]),
                coefs=np.ascontiguousarray(z["gcoefs"]))


# ----------------------------------------------------------------- golden
def test_golden_forces_vacuum(orc):
    z = np.load(os.path.join(GOLD, "forces_vacuum.npz"))
    for k in range(int(z["n"])):
        m = _mol(z, k)
        f, fng, fngpi, E = orc.eval(m)
        scale = max(1.0, np.abs(z["m%d_f" % k]).max())
        assert np.abs(f - z["m%d_f" % k]).max() < 1e-11 * scale
        assert np.abs(fng - z["m%d_fng" % k]).max() < 1e-11 * scale
        assert np.abs(fngpi - z["m%d_fngpi" % k]).max() < 1e-11 * scale
        assert abs(E - float(z["m%d_E" % k])) < 1e-11 * max(1.0, abs(E))


def test_golden_forces_grid(orc):
    z = np.load(os.path.join(GOLD, "forces_grid.npz"))
    g = _grid(z)
    for k in range(int(z["n"])):
        m = _mol(z, k)
        f, _, _, E = orc.eval(m, grid=g)
        assert np.abs(f - z["m%d_f" % k]).max() < 1e-11 * max(1.0, np.abs(f).max())
        assert abs(E - float(z["m%d_E" % k])) < 1e-11 * max(1.0, abs(E))


def test_golden_grid_samples(orc):
    z = np.load(os.path.join(GOLD, "grid_samples.npz"))
    g = _grid(z)
    for p, q, E, F in zip(z["P"], z["PLQ"], z["E"], z["F"]):
        e, f = orc.grid_sample(g, p, q)
        assert abs(e - E) < 1e-11 * max(1.0, abs(E))
        assert np.abs(f - F).max() < 1e-11 * max(1.0, np.abs(F).max())


def test_golden_backneighs(orc):
    z = np.load(os.path.join(GOLD, "backneighs.npz"))
    for k in range(int(z["n"])):
        ng = z["m%d_neighs" % k]
        nnode = int(z["m%d_nnode" % k])
        r, bk, ng_out = orc.make_back_neighs(len(ng), nnode, ng)
        assert r == 0
        assert (bk == z["m%d_bk" % k]).all()
        assert (ng_out == z["m%d_neighs_out" % k]).all()
        r2, bk2 = native.make_back_neighs(len(ng), nnode, ng)  # product topology, bit-exact
        assert r2 == 0 and (bk2 == z["m%d_bk" % k]).all()


def test_golden_fire_trajectory(orc):
    z = np.load(os.path.join(GOLD, "fire_traj.npz"))
    m = {k: z[k] for k in ("apos", "REQs", "neighs", "apars", "bLs", "bKs", "Ksp", "Kpp")}
    m["natoms"], m["nnode"] = int(z["natoms"]), int(z["nnode"])
    n0, apos0, E0, traj = orc.relax(m, FIRE_REF, 1e-4, 5000, 0, ntraj=40)
    assert np.abs(traj - z["traj"]).max() < 1e-12
    assert n0 == int(z["nsteps_fmax"])
    assert abs(E0 - float(z["E_fmax"])) < 1e-10
    n1, _, E1, _ = orc.relax(m, FIRE_REF, 1e-4, 5000, 1)
    assert n1 == int(z["nsteps_ref"])
    assert abs(E1 - float(z["E_ref"])) < 1e-10


def test_golden_terms(orc):
    z = np.load(os.path.join(GOLD, "terms.npz"))
    for kind in range(3):
        E, F = orc.sample_terms(kind, 1.7, 2.0, 1.1, 1.3, z["angles"])
        assert np.abs(E - z["E%d" % kind]).max() < 1e-13
        assert np.abs(F - z["F%d" % kind]).max() < 1e-13
    E, F = orc.sample_pair(z["rs"], z["REQij"])
    assert np.abs(E - z["Epair"]).max() < 1e-12 * np.abs(z["Epair"]).max()
    assert np.abs(F - z["Fpair"]).max() < 1e-12 * np.abs(z["Fpair"]).max()


# ------------------------------------------------- live reference build
@pytest.mark.parametrize("n,seed", [(16, 31), (33, 32), (90, 33), (200, 34)])
def test_oracle_matches_reference_build(orc, ref, n, seed):
    m = native.gen_molecule(n, seed, 0.1, seed + 7)
    fo, ngo, npo, Eo = orc.eval(m)
    fr, ngr, npr, Er = ref.eval(m)
    sc = max(1.0, np.abs(fr).max())
    assert np.abs(fo - fr).max() < 1e-11 * sc
    assert np.abs(ngo - ngr).max() < 1e-11 * sc
    assert np.abs(npo - npr).max() < 1e-11 * sc
    assert abs(Eo - Er) < 1e-11 * max(1.0, abs(Er))


@pytest.mark.parametrize("sw", [(1, 1, 1, 0, 1), (1, 0, 1, 1, 1), (1, 1, 0, 1, 1), (1, 1, 1, 1, 0), (0, 1, 1, 1, 1)])
def test_oracle_switches_match_reference(orc, ref, sw):
    m = native.gen_molecule(48, 5, 0.1, 6)
    fo, _, _, Eo = orc.eval(m, sw=sw)
    fr, _, _, Er = ref.eval(m, sw=sw)
    assert np.abs(fo - fr).max() < 1e-11 * max(1.0, np.abs(fr).max())
    assert abs(Eo - Er) < 1e-11 * max(1.0, abs(Er))


def test_oracle_grid_matches_reference_build(orc, ref, nacl_grid):
    g = nacl_grid
    m = native.gen_workload_system(1, 5, 10001, 16, 256, 0.1, g)
    fo, _, _, Eo = orc.eval(m, grid=g)
    fr, _, _, Er = ref.eval(m, grid=g)
    assert np.abs(fo - fr).max() < 1e-11 * max(1.0, np.abs(fr).max())
    assert abs(Eo - Er) < 1e-11 * max(1.0, abs(Er))
    rng = np.random.default_rng(3)
    for _ in range(50):
        p = np.array([rng.uniform(-9, 9), rng.uniform(-9, 9), rng.uniform(-1.0, 21.0)])  # incl. z clamp range
        q = np.array([rng.uniform(1, 50), rng.uniform(0.1, 3), rng.uniform(-1, 1), 0.0])
        e1, f1 = orc.grid_sample(g, p, q)
        e2, f2 = ref.grid_sample(g, p, q)
        assert abs(e1 - e2) < 1e-10 * max(1.0, abs(e2))
        assert np.abs(f1 - f2).max() < 1e-10 * max(1.0, np.abs(f2).max())


@pytest.mark.parametrize("conv_mode", [0, 1])
def test_oracle_fire_matches_reference_build(orc, ref, conv_mode):
    m = native.gen_molecule(30, 9, 0.1, 10)
    n1, a1, E1, t1 = orc.relax(m, FIRE_REF, 1e-3, 3000, conv_mode, ntraj=60)
    n2, a2, E2, t2 = ref.relax(m, FIRE_REF, 1e-3, 3000, conv_mode, ntraj=60)
    assert n1 == n2
    assert np.abs(t1 - t2).max() < 1e-11
    assert abs(E1 - E2) < 1e-9


def test_backneighs_match_reference_build(ref):
    for n, s in [(20, 1), (100, 2), (256, 3)]:
        m = native.gen_molecule(n, s)
        r, bk_ref, ng_ref = ref.make_back_neighs(n, m["nnode"], m["neighs"])
        r2, bk = native.make_back_neighs(n, m["nnode"], m["neighs"])
        assert (bk == bk_ref).all()
        assert (ng_ref == m["neighs"]).all()  # generator's cap neighbours follow makeBackNeighs
