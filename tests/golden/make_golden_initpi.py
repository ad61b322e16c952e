"""Writes tests/golden/initpi.npz from the reference's own
MMFFsp3_loc::initPi + flipPis (oracle/_ref/libfcref.so, built from
/root/reference by oracle/Makefile).  Run where the reference exists:

    make -C oracle all ref && python tests/golden/make_golden_initpi.py

Inputs: 6 synthetic molecules (16-150 atoms) with perturbed geometry, their
neighbour lists and Ksp/Kpp; plus two hand-made cases (planar sp2 node with
three pi-sigma neighbours; a node inheriting its pi-vector through Kpp).
Outputs: the reference's pi-vectors (nnode x 3) with and without flipPis.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402
from firecore_amd import native  # noqa: E402


def cases():
    out = []
    for k, (n, s) in enumerate([(16, 11), (24, 12), (40, 13), (77, 14), (120, 15), (150, 16)]):
        m = native.gen_molecule(n, s, 0.2, s + 500)
        N = m["natoms"]
        out.append((m["apos"][:N].copy(), m["neighs"], m["nnode"], m["Ksp"], m["Kpp"]))
    # planar node 0 with three pi-sigma neighbours (1, 2, 3), node 1 bonded by Kpp only
    apos = np.array([[0, 0, 0], [1.4, 0.1, 0.05], [-0.7, 1.2, -0.1], [-0.7, -1.2, 0.1], [2.5, 0.2, 0.9]], float)
    neighs = np.array([[1, 2, 3, -1], [0, 4, -1, -1], [0, -1, -1, -1], [0, -1, -1, -1], [1, -1, -1, -1]], np.int32)
    Ksp = np.array([[0.5, 0.5, 0.5, 0], [0, 0, 0, 0]], float)
    Kpp = np.array([[0.3, 0, 0, 0], [0.3, 0, 0, 0]], float)
    out.append((apos, neighs, 2, Ksp, Kpp))
    return out


def main():
    z = {}
    cs = cases()
    for k, (apos, neighs, nnode, Ksp, Kpp) in enumerate(cs):
        z["c%d_apos" % k], z["c%d_neighs" % k], z["c%d_nnode" % k] = apos, neighs, np.int32(nnode)
        z["c%d_Ksp" % k], z["c%d_Kpp" % k] = Ksp, Kpp
        z["c%d_pi" % k] = oracle.ref_init_pi(apos, neighs, nnode, Ksp, Kpp, flip=False)
        z["c%d_pi_flip" % k] = oracle.ref_init_pi(apos, neighs, nnode, Ksp, Kpp, flip=True)
    np.savez_compressed(os.path.join(HERE, "initpi.npz"), n=len(cs), **z)
    print("wrote initpi.npz", len(cs), "cases")


if __name__ == "__main__":
    main()
