"""Periodic molecules (MMFFsp3_loc with bPBC).

Reference: image shifts NBFF::makePBCshifts / MolWorld_sp3::evalPBCshifts
(NBFF.h:42-56, MolWorld_sp3.h:756-765), bonded-neighbour images
MMFFsp3_loc::makeNeighCells(npbc, shifts) (MMFFsp3_loc.h:1394-1434), bonds at
their image in eval_atom (MMFFsp3_loc.h:325-329) and the image sum of
NBFF::evalLJQs_ng4_PBC_atom_omp (NBFF.h:575-620).

CPU: the product's host code (shifts, neighbour images) is bit-exact against
the reference build; the oracle's periodic forces and energy are pinned
against the reference build (1e-10 relative) and against the invariance under
moving atoms by lattice vectors (which changes which image each bond uses).
GPU: forces/energies of periodic molecules from the device against the FP64
oracle (max|dF| <= 1e-4 + 5e-6 max|F| eV/A, |dE| <= 1e-4 + 1e-6 |E| eV), and
relaxations confirmed by the oracle at the device geometry (max|F| < 1.01e-3
eV/A, |dE| < 1e-4 eV), in a batch mixed with vacuum molecules.
"""
import numpy as np
import pytest

from conftest import FIRE_MULTI


def P(*a):
    print(*a, flush=True)


def periodic_molecule(natoms, seed, cell=(11.0, 12.0, 25.0), wrap=True):
    """A synthetic molecule in a box whose atoms with x > 0 (and, separately,
    y > 0 for the second half) are moved back by one lattice vector, so that
    bonds cross the cell boundary and use non-central images."""
    from firecore_amd import native
    m = native.gen_molecule(natoms, seed, 0.05, seed)
    N = m["natoms"]
    apos = np.array(m["apos"])
    apos[:N] -= apos[:N].mean(0)
    lvec = np.diag(np.asarray(cell, float))
    if wrap:
        for i in range(N):
            if apos[i, 0] > 0.5:
                apos[i] -= lvec[0]
            if apos[i, 1] > 0.8 and i % 2 == 0:
                apos[i] -= lvec[1]
    m["apos"] = apos
    return m, lvec


@pytest.mark.parametrize("nPBC", [(1, 1, 0), (1, 1, 1), (2, 1, 0)])
def test_neigh_cells_match_reference(ref, nPBC):
    from firecore_amd import native
    m, lvec = periodic_molecule(40, 3)
    sh, nc = native.make_neigh_cells(m["apos"][:m["natoms"]], m["neighs"], lvec, nPBC)
    sh_r, nc_r = ref.neigh_cells_ref(m, lvec, nPBC)
    assert np.array_equal(nc, nc_r)
    assert np.array_equal(sh, sh_r)
    assert (nc[m["neighs"] >= 0] >= 0).all()
    assert (nc != (len(sh) - 1) // 2).any()  # some bonds cross the cell boundary


def _set_pbc(orc, m, lvec, nPBC):
    from firecore_amd import native
    sh, nc = native.make_neigh_cells(m["apos"][:m["natoms"]], m["neighs"], lvec, nPBC)
    orc.set_pbc(m["natoms"], sh, nc)


@pytest.mark.parametrize("seed,n", [(3, 40), (5, 64)])
def test_oracle_pbc_forces_match_reference(orc, ref, seed, n):
    m, lvec = periodic_molecule(n, seed)
    nPBC = (1, 1, 0)
    _set_pbc(orc, m, lvec, nPBC)
    ref.set_pbc_ref(lvec, nPBC)
    try:
        fo, fng, fngpi, Eo = orc.eval(m)
        fr, rng, rngpi, Er = ref.eval(m)
    finally:
        orc.set_pbc()
        ref.set_pbc_ref()
    sc = np.abs(fr).max()
    assert np.abs(fo - fr).max() < 1e-10 * sc, np.abs(fo - fr).max()
    assert np.abs(fng - rng).max() < 1e-10 * sc
    assert abs(Eo - Er) < 1e-10 * (1 + abs(Er))
    # the image sum matters: the vacuum forces differ
    fv, _, _, Ev = orc.eval(periodic_molecule(n, seed, wrap=False)[0])
    assert abs(Ev - Eo) > 1e-6


def test_oracle_pbc_invariant_under_lattice_moves(orc):
    """Moving atoms by lattice vectors changes which image each bond uses but
    not the physics: forces and energy stay the same.  The image sum of the
    reference is truncated at -nPBC..nPBC cells around each partner, so this
    holds up to the interactions at the edge of that range: the molecule is
    made neutral (no 1/r tail) and three image shells are summed, which leaves
    LJ terms at >= 30 A (measured: 5e-5, 1.4e-6, 1.6e-7 eV/A for one, two,
    three shells)."""
    a, lvec = periodic_molecule(48, 7, wrap=False)
    b, _ = periodic_molecule(48, 7, wrap=True)
    for m in (a, b):
        m["REQs"] = np.array(m["REQs"])
        m["REQs"][:, 2] = 0.0
    res = []
    for m in (a, b):
        _set_pbc(orc, m, lvec, (3, 3, 0))
        try:
            res.append(orc.eval(m))
        finally:
            orc.set_pbc()
    (fa, _, _, Ea), (fb, _, _, Eb) = res
    assert not np.allclose(a["apos"], b["apos"])
    assert np.abs(fa - fb).max() < 1e-6, np.abs(fa - fb).max()
    assert abs(Ea - Eb) < 1e-5, (Ea, Eb)


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_pbc_forces_match_oracle(orc):
    from firecore_amd import native
    mols, cells = [], []
    for n, s in [(40, 3), (64, 5), (90, 9)]:
        m, lvec = periodic_molecule(n, s)
        mols.append(m)
        cells.append(lvec)
    vac = native.gen_molecule(30, 11, 0.1, 11)
    native.init_systems(mols + [vac])
    nPBC = (1, 1, 0)
    for s, lvec in enumerate(cells):
        native.set_lattice(s, lvec, nPBC)
    P("case: periodic molecules, forces")
    native.eval_forces()
    B = native.Batch()
    for s in range(B.nsys):
        m = B.system(s)
        if s < len(cells):
            _set_pbc(orc, m, cells[s], nPBC)
        try:
            fo, _, _, Eo = orc.eval(m)
        finally:
            orc.set_pbc()
        d = np.abs(B.forces(s) - fo).max()
        assert d <= 1e-4 + 5e-6 * np.abs(fo).max(), (s, d)
        assert abs(B.E[s] - Eo) <= 1e-4 + 1e-6 * abs(Eo), (s, B.E[s], Eo)


@pytest.mark.gpu
def test_pbc_relax_mixed_batch(orc):
    """Periodic molecules (split kernels), vacuum molecules (persistent kernel)
    and one 300-atom molecule (split fallback) relaxed in one batch."""
    from firecore_amd import native
    mols, cells = [], []
    for n, s in [(40, 3), (64, 5)]:
        m, lvec = periodic_molecule(n, s)
        mols.append(m)
        cells.append(lvec)
    others = [native.gen_molecule(n, s, 0.1, s) for n, s in [(30, 21), (120, 22), (300, 23)]]
    native.set_opt(*FIRE_MULTI)
    native.init_systems(mols + others)
    nPBC = (1, 1, 0)
    for s, lvec in enumerate(cells):
        native.set_lattice(s, lvec, nPBC)
    try:
        P("case: mixed periodic / vacuum / oversize relax")
        native.relax(20000, 1e-3, 0, True, True)
        B = native.Batch()
        assert (np.array(B.status) == native.ST_CONVERGED).all(), np.array(B.status)
        for s, m0 in enumerate(mols + others):
            m = dict(m0, apos=B.positions(s))  # exact double inputs at the device geometry
            if s < len(cells):
                _set_pbc(orc, m, cells[s], nPBC)
            try:
                fo, _, _, Eo = orc.eval(m)
            finally:
                orc.set_pbc()
            fm = np.sqrt((fo ** 2).sum(1)).max()
            assert fm < 1.01e-3, (s, fm)
            assert abs(B.E[s] - Eo) < 1e-4, (s, B.E[s], Eo)
    finally:
        native.set_opt()
