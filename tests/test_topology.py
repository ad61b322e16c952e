"""Topology, exclusions and the synthetic generator (host code, bit-exact)."""
import numpy as np
import pytest

from firecore_amd import native


def brute_weights(natoms, nnode, neighs, subtract13=True):
    """w(i,j) = [j not in neighs[i]] - #{nodes k: i,j both neighbours of k}."""
    w = np.ones((natoms, natoms), int)
    np.fill_diagonal(w, 0)
    for i in range(natoms):
        for j in neighs[i]:
            if j >= 0:
                w[i, j] = 0
    if subtract13:
        for k in range(nnode):
            ng = [x for x in neighs[k] if x >= 0]
            for a in range(len(ng)):
                for b in range(a + 1, len(ng)):
                    w[ng[a], ng[b]] -= 1
                    w[ng[b], ng[a]] -= 1
    return w


@pytest.mark.parametrize("n,seed", [(16, 1), (60, 2), (256, 3)])
def test_exclusions_match_definition(n, seed):
    m = native.gen_molecule(n, seed)
    bits, ncorr = native.make_exclusions(n, m["nnode"], m["neighs"], True)
    w = brute_weights(n, m["nnode"], m["neighs"])
    assert ((w != 1) == bits).all()
    assert ncorr == int(((w != 1) & (w != 0)).sum())
    assert ncorr == 0  # trees have no ring corrections
    bits2, _ = native.make_exclusions(n, m["nnode"], m["neighs"], False)
    assert ((brute_weights(n, m["nnode"], m["neighs"], False) != 1) == bits2).all()


def test_ring_exclusion_weights():
    """A 4-ring C0-C1-C2-C3 with H caps: the 1-3 pairs (0,2),(1,3) are reached
    through two nodes and get weight -1 (subtracted twice by the reference)."""
    nnode = 4
    neighs = [[1, 3, 4, 5], [0, 2, 6, 7], [1, 3, 8, 9], [2, 0, 10, 11]] + [[k // 2 - 2, -1, -1, -1] for k in range(4, 12)]
    neighs = np.array(neighs, np.int32)
    bits, ncorr = native.make_exclusions(12, nnode, neighs, True)
    w = brute_weights(12, nnode, neighs)
    assert w[0, 2] == -1 and w[1, 3] == -1
    assert ((w != 1) == bits).all()
    assert ncorr == int(((w != 1) & (w != 0)).sum()) == 4


def test_back_neighbours_invert_neighbours():
    for n, s in [(30, 5), (150, 6)]:
        m = native.gen_molecule(n, s)
        r, bk = native.make_back_neighs(n, m["nnode"], m["neighs"])
        assert r == 0
        seen = 0
        for j in range(n):
            for b in bk[j]:
                if b < 0:
                    continue
                ia, k = divmod(int(b), 4)
                assert m["neighs"][ia, k] == j and ia < m["nnode"]
                seen += 1
        assert seen == int((m["neighs"][:m["nnode"]] >= 0).sum())
        # caps have exactly one back-neighbour (their node)
        for c in range(m["nnode"], n):
            assert (bk[c] >= 0).sum() == 1 and bk[c, 0] // 4 == m["neighs"][c, 0]


@pytest.mark.parametrize("n", [5, 16, 24, 100, 256])
def test_generator_invariants(n):
    m = native.gen_molecule(n, 42, 0.1, 43)
    N, M = m["natoms"], m["nnode"]
    assert N == n and 0 < M < N
    ng = m["neighs"]
    # symmetric bonds, caps bonded to one node, nodes first
    for i in range(M):
        for j in ng[i]:
            if j >= 0:
                assert i in ng[j]
    for c in range(M, N):
        assert ng[c, 0] < M and (ng[c, 1:] == -1).all()
    assert abs(m["REQs"][:, 2].sum()) < 1e-12
    assert np.allclose(np.linalg.norm(m["apos"][N:], axis=1), 1.0)
    assert set(np.unique(m["elem"])) <= {0, 1, 2, 3}
    assert ((m["elem"][M:]) == 3).all()
    # parameter ranges of the workload definition
    assert (m["REQs"][:, 0] >= 1.2).all() and (m["REQs"][:, 0] <= 1.95).all()
    E0 = m["REQs"][:, 1] ** 2
    assert (E0 >= 0.002 - 1e-12).all() and (E0 <= 0.01).all()
    bl = m["bLs"][m["bLs"] > 0]
    assert bl.min() >= 1.09 - 1e-12 and bl.max() <= 1.50 + 1e-12


def test_generator_deterministic():
    a = native.gen_molecule(77, 9, 0.1, 10)
    b = native.gen_molecule(77, 9, 0.1, 10)
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert (a[k] == b[k]).all()
    c = native.gen_molecule(77, 9, 0.1, 11)
    assert not (c["apos"] == a["apos"]).all()


def test_workload_systems(nacl_grid):
    sizes = []
    for i in range(20):
        m = native.gen_workload_system(1, i, 10001, 16, 256, 0.1, nacl_grid)
        sizes.append(m["natoms"])
        zmin = m["apos"][:m["natoms"], 2].min()
        assert 2.7 - 1e-9 <= zmin - nacl_grid["pos0"][2] <= 3.3 + 1e-9
    assert min(sizes) >= 16 and max(sizes) <= 256
    p = native.gen_workload_system(2, 3, 20001, 8, 64, 0.1, nacl_grid)
    assert 16 <= p["natoms"] <= 128
