"""Initial pi-vectors (init_pi_vectors, restating MMFFsp3_loc::initPi +
flipPis, MMFFsp3_loc.h:882-945 / :1167-1172) pinned three ways, CPU only:
the golden fixture written from the reference's own build, that build itself
where it is present, and the geometry (unit normal of the sigma-bond plane,
inheritance through the strongest pi-pi bond, +z otherwise, flip half-space).
"""
import os

import numpy as np
import pytest

import oracle
from firecore_amd import native

GOLD = os.path.join(os.path.dirname(__file__), "golden", "initpi.npz")


def test_golden_initpi():
    z = np.load(GOLD)
    for k in range(int(z["n"])):
        args = (z["c%d_apos" % k], z["c%d_neighs" % k], int(z["c%d_nnode" % k]), z["c%d_Ksp" % k], z["c%d_Kpp" % k])
        for flip, key in ((False, "pi"), (True, "pi_flip")):
            got = native.init_pi_vectors(*args, flip=flip)
            assert np.abs(got - z["c%d_%s" % (k, key)]).max() < 1e-14, (k, key)


def test_initpi_matches_reference_build():
    for s in range(5):
        m = native.gen_molecule(30 + 40 * s, 100 + s, 0.15, 900 + s)
        N, M = m["natoms"], m["nnode"]
        ref = oracle.ref_init_pi(m["apos"][:N], m["neighs"], M, m["Ksp"], m["Kpp"])
        if ref is None:
            pytest.skip("reference build oracle/_ref is not present on this machine")
        got = native.init_pi_vectors(m["apos"][:N], m["neighs"], M, m["Ksp"], m["Kpp"])
        assert np.abs(got - ref).max() < 1e-14, s


def test_initpi_geometry():
    rng = np.random.default_rng(3)
    # node 0: three pi-sigma neighbours in general position -> normal of their plane
    apos = np.array([[0, 0, 0], [1.4, 0.1, 0.05], [-0.7, 1.2, -0.1], [-0.7, -1.2, 0.1], [2.5, 0.2, 0.9],
                     [0.0, 0.0, 3.0]], float)
    apos += rng.normal(scale=0.05, size=apos.shape)
    neighs = np.array([[1, 2, 3, -1], [0, 4, -1, -1], [0, -1, -1, -1], [0, -1, -1, -1], [1, -1, -1, -1],
                       [-1, -1, -1, -1]], np.int32)
    # node 1: two pi-sigma neighbours (0 and 4) -> normal of the plane (0, 4, node 1)
    # node 2: no pi-sigma neighbour, Kpp bond to node 0 -> copies node 0's pi
    # node 3: nothing -> +z
    Ksp = np.array([[0.5, 0.5, 0.5, 0], [0.4, 0.4, 0, 0], [0, 0, 0, 0], [0, 0, 0, 0]], float)
    Kpp = np.array([[0, 0, 0, 0], [0, 0, 0, 0], [0.25, 0, 0, 0], [0, 0, 0, 0]], float)
    neighs[2] = [0, -1, -1, -1]
    pi = native.init_pi_vectors(apos, neighs, 4, Ksp, Kpp, flip=False)
    n0 = np.cross(apos[1] - apos[3], apos[2] - apos[3])
    assert np.allclose(pi[0], n0 / np.linalg.norm(n0), atol=1e-14)
    n1 = np.cross(apos[0] - apos[1], apos[4] - apos[1])
    assert np.allclose(pi[1], n1 / np.linalg.norm(n1), atol=1e-14)
    assert np.array_equal(pi[2], pi[0])
    assert np.array_equal(pi[3], [0.0, 0.0, 1.0])
    # the pi-vector of node 0 is perpendicular to the plane of its three neighbours
    for a, b in ((1, 2), (2, 3), (1, 3)):
        assert abs(np.dot(pi[0], apos[a] - apos[b])) < 1e-12
    flipped = native.init_pi_vectors(apos, neighs, 4, Ksp, Kpp, flip=True)
    assert (flipped.sum(1) >= 0).all()
    assert np.allclose(np.abs(flipped), np.abs(pi))
