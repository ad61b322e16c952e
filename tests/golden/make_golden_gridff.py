"""Writes tests/golden/gridff_points.npz: GridFF node potentials (Pauli,
London, bare Coulomb) of the synthetic NaCl(001) slab at 160 points, from the
reference's own GridFF::evalGridFFPoint_Mors / evalGridFFPoint_Coul
(oracle/_ref/libfcref.so, built from /root/reference by oracle/Makefile) with
the image ranges of its B-spline grid generation (Morse: autoNPBC with 20 A,
(3,3) for the 5.64 A cell; Coulomb: (25,25), GridFF.h:849).  Run where the
reference exists:

    make -C oracle all ref && python tests/golden/make_golden_gridff.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

A = 5.64


def nacl_slab(a=A, q=0.9):
    """The synthetic substrate: conventional (001) cell, 4 ions per layer,
    2 layers, top layer at z = 0 (Na: R0 1.4915 A, E0 0.00130092313 eV;
    Cl: R0 1.9735 A, E0 0.00984365166 eV)."""
    RNa, ENa, RCl, ECl = 1.4915, 0.00130092313, 1.9735, 0.00984365166
    apos, REQ = [], []
    for L in range(2):
        z = -0.5 * a * L
        for k, (x, y) in enumerate([(0, 0), (0.5, 0.5), (0.5, 0), (0, 0.5)]):
            na = (k < 2) ^ bool(L & 1)
            apos.append([x * a, y * a, z])
            REQ.append([RNa if na else RCl, np.sqrt(ENa if na else ECl), q if na else -q, 0.0])
    return np.array(apos), np.array(REQ), np.array([a, 0, 0, 0, a, 0, 0, 0, 20.0])


def main():
    apos, REQ, lvec = nacl_slab()
    rng = np.random.default_rng(42)
    n = 160
    pts = np.stack([rng.uniform(-A, A, n), rng.uniform(-A, A, n), rng.uniform(0.2, 12.0, n)], 1)
    ref = oracle.ref_gridff_points(apos, REQ, lvec, 1.5, (3, 3), (25, 25), pts)
    assert ref is not None, "build oracle/_ref first (make -C oracle ref)"
    np.savez_compressed(os.path.join(HERE, "gridff_points.npz"), apos=apos, REQs=REQ, lvec=lvec, alpha=1.5,
                        pts=pts, PLQ=ref)
    print("wrote gridff_points.npz", n, "points")


if __name__ == "__main__":
    main()
