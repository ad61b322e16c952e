"""Explicit substrate atoms instead of the GridFF (MolWorld_sp3::loadSurf with
bGrid = false; per molecule atom GridFF::evalMorsePBC, GridFF.h:1229-1261,
with addAtomicForceMorseQ, Forces.h:540-561, and combineREQ mixing).

CPU: the oracle's substrate term is pinned against the reference build
(1e-10 relative) and by central finite differences of its energy.
GPU: device forces and energies over explicit NaCl atoms against the FP64
oracle (max|dF| <= 1e-4 + 5e-6 max|F| eV/A, |dE| <= 1e-4 + 1e-6 |E| eV), and
relaxations confirmed by the oracle at the device geometry (max|F| < 1.01e-3
eV/A, |dE| < 1e-4 eV with the exact double inputs).
"""
import numpy as np
import pytest

from conftest import FIRE_MULTI

NPBC = (2, 2, 0)


def P(*a):
    print(*a, flush=True)


def slab():
    from firecore_amd import native
    a, q, lv, _ = native.surface_atoms("synth:NaCl")
    return a, q, lv


def molecule_over_slab(n, seed, z=3.0):
    from firecore_amd import native
    m = native.gen_molecule(n, seed, 0.1, seed)
    N = m["natoms"]
    apos = np.array(m["apos"])
    apos[:N] -= apos[:N].mean(0)
    apos[:N, 2] += z - apos[:N, 2].min()
    m["apos"] = apos
    return m


@pytest.mark.parametrize("seed,n", [(3, 30), (8, 60)])
def test_oracle_surface_matches_reference(orc, ref, seed, n):
    a, q, lv = slab()
    m = molecule_over_slab(n, seed)
    orc.set_surface(a, q, lv, NPBC, 1.5)
    ref.set_surface(a, q, lv, NPBC, 1.5)
    try:
        fo, _, _, Eo = orc.eval(m)
        fr, _, _, Er = ref.eval(m)
    finally:
        orc.set_surface()
        ref.set_surface()
    sc = np.abs(fr).max()
    assert np.abs(fo - fr).max() < 1e-10 * sc, np.abs(fo - fr).max()
    assert abs(Eo - Er) < 1e-10 * (1 + abs(Er))
    f0, _, _, E0 = orc.eval(m)  # without the substrate
    assert abs(E0 - Eo) > 1e-3


def test_oracle_surface_force_is_energy_gradient(orc):
    """F = -dE/dx of the substrate term by central differences (h = 1e-3 A;
    the reference rounds each pair's Morse amplitude to float, which makes the
    energy step by ~1e-8 relative, so h stays large): pins the sign, the
    factor 2 of the Morse force and the Coulomb 1/r^3."""
    a, q, lv = slab()
    m = molecule_over_slab(24, 5, z=2.6)
    sw = (0, 0, 0, 0, 0)  # molecule terms off: only the substrate acts
    orc.set_surface(a, q, lv, NPBC, 1.5)
    try:
        f, _, _, _ = orc.eval(m, sw=sw)
        rng = np.random.default_rng(3)
        h = 1e-3
        for ia in rng.choice(m["natoms"], 6, replace=False):
            for d in range(3):
                p1 = np.array(m["apos"])
                p2 = np.array(m["apos"])
                p1[ia, d] += h
                p2[ia, d] -= h
                E1 = orc.eval(m, apos=p1, sw=sw)[3]
                E2 = orc.eval(m, apos=p2, sw=sw)[3]
                fd = -(E1 - E2) / (2 * h)
                assert abs(fd - f[ia, d]) < 1e-5 + 1e-4 * abs(f[ia, d]), (ia, d, fd, f[ia, d])
    finally:
        orc.set_surface()


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_surface_atoms_forces_match_oracle(orc):
    from firecore_amd import native
    mols = [molecule_over_slab(n, s, z) for n, s, z in [(20, 31, 2.8), (48, 32, 3.0), (90, 33, 3.3)]]
    native.init_systems(mols)
    native.attach_surface_atoms("synth:NaCl", NPBC)
    a, q, lv, npbc = native.surface_atoms()
    assert npbc == NPBC and len(a) == 8
    P("case: explicit substrate atoms, forces")
    try:
        native.eval_forces()
        B = native.Batch()
        orc.set_surface(a, q, lv, NPBC, 1.5)
        for s in range(B.nsys):
            m = B.system(s)
            fo, _, _, Eo = orc.eval(m)
            d = np.abs(B.forces(s) - fo).max()
            assert d <= 1e-4 + 5e-6 * np.abs(fo).max(), (s, d)
            assert abs(B.E[s] - Eo) <= 1e-4 + 1e-6 * abs(Eo), (s, B.E[s], Eo)
    finally:
        orc.set_surface()
        native.attach_surface_atoms(None)


@pytest.mark.gpu
def test_surface_atoms_relax(orc):
    """Relaxation over the explicit substrate.  The molecules are neutral: the
    substrate term carries the reference's bare Coulomb, under which a random
    charged atom collapses onto an ion of opposite charge (the FP64 oracle
    reaches E ~ -1e3 eV for the charged version of the first system); the
    charged force field is covered by the force test above."""
    from firecore_amd import native
    mols = [molecule_over_slab(n, s, z) for n, s, z in [(24, 41, 3.0), (60, 42, 3.2), (40, 43, 3.0)]]
    for m in mols:
        m["REQs"] = np.array(m["REQs"])
        m["REQs"][:, 2] = 0.0
    native.set_opt(*FIRE_MULTI)
    native.set_force_limit(10.0)
    native.init_systems(mols)
    native.attach_surface_atoms("synth:NaCl", NPBC)
    a, q, lv, _ = native.surface_atoms()
    try:
        P("case: explicit substrate atoms, relax")
        native.relax(20000, 1e-3, 0, True, True)
        B = native.Batch()
        assert (np.array(B.status) == native.ST_CONVERGED).all(), np.array(B.status)
        orc.set_surface(a, q, lv, NPBC, 1.5)
        for s, m0 in enumerate(mols):
            m = dict(m0, apos=B.positions(s))
            fo, _, _, Eo = orc.eval(m)
            fm = np.sqrt((fo ** 2).sum(1)).max()
            assert fm < 1.01e-3, (s, fm)
            assert abs(B.E[s] - Eo) < 1e-4, (s, B.E[s], Eo)
    finally:
        orc.set_surface()
        native.attach_surface_atoms(None)
        native.set_force_limit(0.0)
        native.set_opt()
