"""Typed molecules from .xyz files: the builder behind init(xyz_name).

Reference: MolWorld_sp3::init with bMMFF and without electron pairs
(MolWorld_sp3.h:309-357): MMFFparams::loadXYZ, autoBonds / autoBondsPBC,
checkNumberOfBonds, autoConfEPi, assignAllBondParams, sortConfAtomsFirst,
assignTypes, toMMFFsp3_loc, flipPis and the pi-vector normalisation of the
init's check evaluation.

CPU: against the reference's own build of 20 molecules (6 periodic; golden
fixtures written by tests/golden/make_golden_builder.py): topology, image
cells and types bit-exact, every float parameter and position within 1e-12
(measured: <= 2.2e-16).  Where the reference build exists, more molecules
are compared live.  Inputs the reference refuses (it ends the process in
checkNumberOfBonds, MMFFBuilder.h:2209-2237) raise here.
GPU: init(xyz_name) replicas of typed molecules (periodic and vacuum) relaxed
through MMFF_multi.run; the device holds the builder's parameters, and the
FP64 oracle at the device geometry confirms the minimum (max|F| < 1.01e-3
eV/A, |dE| < 1e-4 eV, exact double parameters).
"""
import os

import numpy as np
import pytest

from conftest import FIRE_MULTI

HERE = os.path.dirname(os.path.abspath(__file__))
TYPED = os.path.join(HERE, "golden", "typed")
GOLD = os.path.join(HERE, "golden", "typed_build.npz")
FLOATS = ("apos", "REQs", "apars", "bLs", "bKs", "Ksp", "Kpp")


def P(*a):
    print(*a, flush=True)


def golden():
    with np.load(GOLD) as z:
        names = [str(n) for n in z["names"]]
        return {n: {k.split("/", 1)[1]: z[k] for k in z.files if k.startswith(n + "/")} for n in names}


GOLDEN = golden()
PERIODIC = [n for n, g in GOLDEN.items() if np.abs(g["lvec"]).max() > 0]


def build(name):
    from firecore_amd import native
    return native.build_xyz(os.path.join(TYPED, name + ".xyz"), TYPED)


@pytest.mark.parametrize("name", list(GOLDEN))
def test_build_matches_reference_golden(name):
    g = GOLDEN[name]
    m, lv = build(name)
    assert (m["natoms"], m["nnode"]) == (int(g["natoms"]), int(g["nnode"]))
    assert np.array_equal(m["neighs"], g["neighs"])
    for k in FLOATS:
        d = np.abs(np.asarray(m[k]) - g[k]).max() if g[k].size else 0.0
        assert d <= 1e-12, (k, d)
    assert (lv is not None) == (name in PERIODIC)
    if lv is not None:
        assert np.array_equal(lv, g["lvec"])


@pytest.mark.parametrize("name", PERIODIC)
def test_build_neigh_cells_match_reference(name):
    """Image cells of the bonded neighbours for init's nPBC = (1,1,0)
    (MMFFsp3_loc::makeNeighCells) from the built molecule."""
    from firecore_amd import native
    g = GOLDEN[name]
    m, lv = build(name)
    _, nc = native.make_neigh_cells(m["apos"][:m["natoms"]], m["neighs"], lv, (1, 1, 0))
    assert np.array_equal(nc, g["neighCell"])
    if name == "polydiacetylene":  # a chain bonded across the cell boundary
        assert (nc[m["neighs"] >= 0] != 4).sum() == 2


def test_build_invariants():
    """What the construction fixes whatever the tables hold: unit pi-vectors,
    apars = (cos, sin) of half the sp3 angle, symmetric bond lists with caps
    bonded to one node, Kpp = sqrt(Kpp_i Kpp_j) symmetric between nodes."""
    for name in ("PTCDA", "guanine", "nHexadecan_dicarboxylic", "N_2edge_assay_1"):
        m, _ = build(name)
        N, M = m["natoms"], m["nnode"]
        pi = m["apos"][N:]
        assert np.abs(np.linalg.norm(pi, axis=1) - 1).max() < 1e-14
        assert np.abs(m["apars"][:, 0] ** 2 + m["apars"][:, 1] ** 2 - 1).max() < 1e-14
        ng = m["neighs"]
        for i in range(N):
            for j in ng[i][ng[i] >= 0]:
                assert i in ng[j]
        assert ((ng[M:] >= 0).sum(1) == 1).all() and (ng[M:, 0] < M).all()
        for i in range(M):
            for k, j in enumerate(ng[i]):
                if 0 <= j < M:
                    kk = list(ng[j]).index(i)
                    assert m["Kpp"][i, k] == m["Kpp"][j, kk]


@pytest.mark.parametrize("name", ["pentacene", "adenine", "thymine", "polymer-2_new", "deoxyglucose"])
def test_build_matches_reference_live(ref, name):
    """The reference build of more molecules from its resource folder."""
    from firecore_amd import native
    res = "/root/reference/cpp"
    r, rl = ref.build_xyz_ref(res, "common_resources/xyz/" + name, (1, 1, 0))
    m, ml = native.build_xyz(res + "/common_resources/xyz/" + name + ".xyz", res + "/common_resources")
    assert (m["natoms"], m["nnode"]) == (r["natoms"], r["nnode"])
    assert np.array_equal(m["neighs"], r["neighs"])
    for k in FLOATS:
        assert np.abs(np.asarray(m[k]) - r[k]).max() <= 1e-12, k
    assert (ml is None) == (rl is None)


@pytest.mark.parametrize("name", PERIODIC)
def test_oracle_matches_reference_on_built_periodic(orc, ref, name):
    """The oracle's periodic force field against the reference's on the built
    molecules; polydiacetylene's 5 A cell makes atoms both 1-2 and 1-3
    partners through different images, and puts 1-3 paths at images outside
    the summed range."""
    from firecore_amd import native
    m, lv = build(name)
    sh, nc = native.make_neigh_cells(m["apos"][:m["natoms"]], m["neighs"], lv, (1, 1, 0))
    orc.set_pbc(m["natoms"], sh, nc)
    ref.set_pbc_ref(lv, (1, 1, 0))
    try:
        fo, _, _, Eo = orc.eval(m)
        fr, _, _, Er = ref.eval(m)
    finally:
        orc.set_pbc()
        ref.set_pbc_ref()
    assert np.abs(fo - fr).max() < 1e-10 * np.abs(fr).max()
    assert abs(Eo - Er) < 1e-10 * (1 + abs(Er))


@pytest.mark.parametrize("text,what", [
    ("2\n\nC 0 0 0 0\nC 10 0 0 0\n", "atom without bonds"),
    ("3\n\nC 0 0 0 0\nH 0.9 0 0 0\nC 1.8 0 0 0\n", "capping atom with two bonds"),
    ("\n3\n\nO 0 0 0 0\nH 0.96 0 0 0\nH -0.24 0.93 0 0\n", "no atom count on the first line"),
])
def test_build_rejects_what_the_reference_refuses(tmp_path, text, what):
    from firecore_amd import native
    f = tmp_path / "bad.xyz"
    f.write_text(text)
    with pytest.raises(ValueError):
        native.build_xyz(str(f), TYPED)
    f.write_text(text.lstrip("\n"))
    if what.startswith("no atom count"):  # the same molecule with its count in place builds
        m, _ = native.build_xyz(str(f), TYPED)
        assert m["natoms"] == 3 and m["nnode"] == 1


def test_build_missing_file_raises():
    from firecore_amd import native
    with pytest.raises(ValueError):
        native.build_xyz(os.path.join(TYPED, "no_such_molecule.xyz"), TYPED)


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_init_xyz_forces_match_oracle(orc):
    """Device forces and energies of the built periodic molecules at their
    file geometry against the FP64 oracle (max|dF| <= 1e-4 + 5e-6 max|F|
    eV/A, |dE| <= 1e-4 + 1e-6 |E| eV)."""
    from firecore_amd import native
    mols, cells = [], []
    for name in PERIODIC:
        m, lv = build(name)
        mols.append(m)
        cells.append(lv)
    native.init_systems(mols)
    for s, lv in enumerate(cells):
        native.set_lattice(s, lv, (1, 1, 0))
    P("case: built periodic molecules, forces", PERIODIC)
    native.eval_forces()
    B = native.Batch()
    for s, (m, lv) in enumerate(zip(mols, cells)):
        sh, nc = native.make_neigh_cells(m["apos"][:m["natoms"]], m["neighs"], lv, (1, 1, 0))
        orc.set_pbc(m["natoms"], sh, nc)
        try:
            fo, _, _, Eo = orc.eval(B.system(s))
        finally:
            orc.set_pbc()
        d = np.abs(B.forces(s) - fo).max()
        assert d <= 1e-4 + 5e-6 * np.abs(fo).max(), (PERIODIC[s], d)
        assert abs(B.E[s] - Eo) <= 1e-4 + 1e-6 * abs(Eo), (PERIODIC[s], B.E[s], Eo)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["PTCDA", "guanine", "nHexadecan_dicarboxylic", "polydiacetylene"])
def test_init_xyz_relax(orc, name):
    from firecore_amd import MMFF_multi as mm
    from firecore_amd import native
    g = GOLDEN[name]
    m0, lv = build(name)
    nsys = 6
    P("case: init(xyz) + run,", name)
    mm.init(nSys_=nsys, xyz_name=os.path.join(TYPED, name), nPBC=(1, 1, 0),
            sElementTypes=os.path.join(TYPED, "ElementTypes.dat"), sAtomTypes=os.path.join(TYPED, "AtomTypes.dat"),
            sBondTypes=os.path.join(TYPED, "BondTypes.dat"))
    B = native.Batch()
    assert B.nsys == nsys
    s0 = B.system(0)
    assert np.array_equal(s0["neighs"], g["neighs"])
    for k in ("REQs", "apars", "bLs", "bKs", "Ksp", "Kpp"):
        assert np.array_equal(s0[k], g[k].astype(np.float32).astype(np.float64)), k
    native.set_opt(*FIRE_MULTI)
    try:
        outE = np.zeros(nsys)
        outF = np.zeros(nsys)
        n = mm.run(20000, Fconv=1e-3, outE=outE, outF=outF)
        assert 0 < n <= 20000 and (outF < 1e-3).all(), outF
        B = native.Batch()
        assert (np.array(B.status) == native.ST_CONVERGED).all(), np.array(B.status)
        if lv is not None:
            sh, nc = native.make_neigh_cells(m0["apos"][:m0["natoms"]], m0["neighs"], lv, (1, 1, 0))
            orc.set_pbc(m0["natoms"], sh, nc)
        try:
            for s in range(nsys):
                m = dict(m0, apos=B.positions(s))
                fo, _, _, Eo = orc.eval(m)
                fm = np.sqrt((fo ** 2).sum(1)).max()
                assert fm < 1.01e-3, (s, fm)
                assert abs(B.E[s] - Eo) < 1e-4, (s, B.E[s], Eo)
                assert abs(outE[s] - B.E[s]) < 1e-9
        finally:
            orc.set_pbc()
    finally:
        native.set_opt()


@pytest.mark.gpu
@pytest.mark.parametrize("path", [0, 1])
def test_typed_trajectory_parity(orc, path):
    """Reference molecules typed by the builder (sp2 pi-systems, N and O lone
    pairs, up to 139 atoms), displaced by 0.1 A from their file geometry: the
    device FIRE trajectory follows the FP64 oracle's over 40 steps within 2e-4 A
    (persistent kernel and split kernels)."""
    from conftest import FIRE_REF
    from firecore_amd import native
    rng = np.random.default_rng(5)
    mols = []
    for name in ("guanine", "Benzene_deriv", "nHexadecan_dicarboxylic", "N_2edge_assay_1"):
        m, _ = build(name)
        m["apos"] = np.array(m["apos"])
        m["apos"][:m["natoms"]] += rng.normal(size=(m["natoms"], 3)) * 0.1
        mols.append(m)
    native.set_force_limit(0.0)
    native.set_opt(*FIRE_REF)
    native.init_systems(mols)
    P("case: typed molecules, 40-step trajectory, path", path)
    try:
        native.relax(40, 1e-12, path, True, True)
        B = native.Batch()
        assert (np.array(B.status) == native.ST_MAXSTEP).all() and (np.array(B.nsteps) == 40).all()
        for s, m in enumerate(mols):
            _, _, _, traj = orc.relax(m, FIRE_REF, 1e-12, 40, 0, ntraj=40)
            d = np.abs(B.positions(s) - traj[-1]).max()
            assert d < 2e-4, (s, d)
            assert np.abs(traj[-1] - m["apos"]).max() > 0.01  # the systems did move
    finally:
        native.set_opt()


@pytest.mark.gpu
def test_periodic_typed_trajectory_parity(orc):
    """Periodic reference molecules (polydiacetylene's chain bonded across the
    cell, PTCDA, BPBA), displaced by 0.05 A: 40 FIRE steps of the split kernels
    against the FP64 oracle with the same image shifts and neighbour images."""
    from conftest import FIRE_REF
    from firecore_amd import native
    rng = np.random.default_rng(6)
    mols, cells = [], []
    for name in ("polydiacetylene", "PTCDA", "BPBA"):
        m, lv = build(name)
        m["apos"] = np.array(m["apos"])
        m["apos"][:m["natoms"]] += rng.normal(size=(m["natoms"], 3)) * 0.05
        mols.append(m)
        cells.append(lv)
    native.set_force_limit(0.0)
    native.set_opt(*FIRE_REF)
    native.init_systems(mols)
    for s, lv in enumerate(cells):
        native.set_lattice(s, lv, (1, 1, 0))
    P("case: periodic typed molecules, 40-step trajectory")
    try:
        native.relax(40, 1e-12, 1, True, True)
        B = native.Batch()
        assert (np.array(B.status) == native.ST_MAXSTEP).all()
        for s, (m, lv) in enumerate(zip(mols, cells)):
            sh, nc = native.make_neigh_cells(m["apos"][:m["natoms"]], m["neighs"], lv, (1, 1, 0))
            orc.set_pbc(m["natoms"], sh, nc)
            try:
                _, _, _, traj = orc.relax(m, FIRE_REF, 1e-12, 40, 0, ntraj=40)
            finally:
                orc.set_pbc()
            d = np.abs(B.positions(s) - traj[-1]).max()
            assert d < 2e-4, (s, d)
    finally:
        native.set_opt()
