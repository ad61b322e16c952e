"""Minima database of the exploring mode (MolecularDatabase::addIfNewDescriptor,
MolecularDatabase.h:430-474, with the default descriptors of setDescriptors,
:377-400; fed in MolWorld_sp3_multi.h:863-879 while bSaveToDatabase is on).

CPU: descriptors (sqrt of the inertia eigenvalues + atom count) and the
superposed nearest-same-type distance against the reference's own
MolecularDatabase (oracle/_ref; its iterative superposition converges to
~1e-6, so 1e-5 relative); the duplicate decisions of a sequence of structures
(rigid motions + small noise = duplicates, large distortions = new members)
identical to the reference's.  Without the reference: rotation/translation
invariance, the closed-form descriptor of a square, type sensitivity and both
sides of the 0.1-per-atom threshold.
GPU: exploring with the database on merges the repeated minima of identical
replicas; every member is a recorded minimum, members are pairwise distinct,
occurrences add up to the minima offered.
"""
import numpy as np
import pytest

from conftest import FIRE_MULTI


def rot(seed):
    from scipy.spatial.transform import Rotation
    return Rotation.random(random_state=seed).as_matrix()


def molecule(n, seed):
    """n atoms on a jittered chain-like cloud, 4 types."""
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.normal(size=(n, 3)), 0) * 1.2
    t = rng.integers(0, 4, n).astype(np.int32)
    return x, t


def moved(x, seed, noise=0.0):
    rng = np.random.default_rng(seed)
    return x @ rot(seed).T + rng.normal(size=3) * 5 + rng.normal(size=x.shape) * noise


# ------------------------------------------------------------------ CPU, against the reference
@pytest.mark.parametrize("n", [4, 12, 40, 150])
def test_descriptor_matches_reference(ref, n):
    from firecore_amd import native
    x, _ = molecule(n, n)
    d = native.database_descriptor(x)
    dr = ref.db_descriptor(x)
    assert len(d) == len(dr) == 4
    assert np.abs(d - dr).max() < 1e-6 * max(1.0, np.abs(dr).max())


@pytest.mark.parametrize("n,noise", [(8, 0.01), (24, 0.05), (60, 0.3), (180, 0.02)])
def test_distance_matches_reference(ref, n, noise):
    from firecore_amd import native
    x, t = molecule(n, 100 + n)
    y = moved(x, n, noise)
    for a, b in ((x, y), (y, x)):
        d = native.database_distance(a, t, b, t)
        dr = ref.db_distance(a, t, b, t)
        assert abs(d - dr) < 1e-5 * max(dr, 1.0), (d, dr)


def test_database_sequence_matches_reference(ref):
    """Three conformations of one molecule, each offered again after rigid
    motions with small noise (duplicates) and with large distortions (new
    members), interleaved: the same decision for every structure."""
    from firecore_amd import native
    base, t = molecule(24, 7)
    rng = np.random.default_rng(3)
    confs = [base, base + rng.normal(size=base.shape) * 0.9, base + rng.normal(size=base.shape) * 0.9]
    seq = []
    for k in range(12):
        c = confs[k % 3]
        seq.append(moved(c, 200 + k, 0.02))
        if k % 4 == 3:
            seq.append(c + rng.normal(size=c.shape) * 0.7)
    native.clear_database()
    ref.db_reset()
    mine = [native.database_add(s, t, -float(i)) for i, s in enumerate(seq)]
    theirs = [ref.db_add(s, t) for s in seq]
    assert mine == theirs, (mine, theirs)
    members = native.database()
    assert len(members) == ref.db_count() == mine.count(-1)
    assert sum(m["occurrence"] for m in members) == len(seq)
    assert 3 <= len(members) < len(seq)


# ------------------------------------------------------------------ CPU, closed forms
def test_descriptor_of_a_square():
    """Four unit-distance atoms at (+-1,0,0), (0,+-1,0): inertia diag(2,2,4)."""
    from firecore_amd import native
    d = native.database_descriptor(np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0.]]))
    assert np.allclose(d, [np.sqrt(2), np.sqrt(2), 2.0, 4.0], atol=1e-14)
    assert len(native.database_descriptor(np.zeros((2, 3)))) == 1  # no principal axes below 3 atoms


def test_distance_invariances_and_types():
    from firecore_amd import native
    x, t = molecule(30, 11)
    y = moved(x, 5)
    assert native.database_distance(x, t, y, t) < 1e-10
    # swapping the positions of two atoms of different types is seen
    t2 = t.copy()
    i, j = 0, int(np.argmax(t != t[0]))
    t2[i], t2[j] = t[j], t[i]
    assert native.database_distance(x, t, x, t2) > 0.5
    # one atom displaced by delta, far from every other atom of its type: distance ~ delta
    d = native.database_distance(x, t, x + np.eye(30, 1).reshape(30, 1) * np.array([0.3, 0, 0]), t)
    assert 0.2 < d < 0.31


def test_threshold_both_sides():
    from firecore_amd import native
    x, t = molecule(20, 21)
    rng = np.random.default_rng(0)
    native.clear_database()
    assert native.database_add(x, t) == -1
    assert native.database_add(moved(x, 1, 0.05), t) == 0        # 0.05 A noise: ~0.22 A total, 0.011 per atom
    assert native.database_add(x + rng.normal(size=x.shape) * 0.8, t) == -1  # ~3.5 A total: > 0.1 per atom
    assert [m["occurrence"] for m in native.database()] == [2, 1]
    native.clear_database()


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_explore_feeds_database():
    from firecore_amd import native
    native.set_opt(*FIRE_MULTI)
    native.set_force_limit(10.0)
    # eight identical replicas of one molecule: their minima repeat
    m = native.gen_molecule(20, 41, 0.0, 41)
    native.init_systems([m] * 8)
    native.clear_database()
    native.set_save_to_database(True)
    kmax = 6
    try:
        print("case: explore with the minima database, 8 replicas of a 20-atom molecule", flush=True)
        found = native.explore(3000, 1e-3, 300.0, 0.1, 150, 0, 7, kmax, True)
        members = native.database()
        offered = sum(min(native.minima(s, kmax)[0], kmax) for s in range(8))
        assert found >= 8 and offered >= 8
        assert sum(mm["occurrence"] for mm in members) == offered
        assert 1 <= len(members) < offered  # the replicas' first minima coincide
        Es = np.concatenate([native.minima(s, kmax)[1] for s in range(8)])
        elem = np.asarray(m["elem"], np.int32)
        for a, ma in enumerate(members):
            assert np.min(np.abs(Es - ma["E"])) < 1e-12
            for mb in members[a + 1:]:
                dd = native.database_descriptor(ma["pos"]) - native.database_descriptor(mb["pos"])
                dist = native.database_distance(ma["pos"], elem, mb["pos"], elem) / ma["natoms"]
                assert np.linalg.norm(dd) >= 5.0 or dist >= 0.1
    finally:
        native.set_save_to_database(False)
        native.clear_database()
        native.set_force_limit(0.0)
        native.set_opt()
