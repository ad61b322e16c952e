"""Writes the golden fixtures of tests/golden from the reference's own CPU
classes (oracle/_ref/libfcref.so, compiled from /root/reference by
oracle/Makefile).  Run in the container where the reference exists:

    make -C oracle all ref && python tests/golden/make_golden.py

Fixtures (inputs and reference outputs only, .npz):
  forces_vacuum.npz   5 molecules: assembled forces, recoils, energies
  forces_grid.npz     2 molecules above a small synthetic NaCl grid
  grid_samples.npz    64 B-spline samples on that grid
  backneighs.npz      back-neighbour lists of the 5 molecules
  fire_traj.npz       first 40 FIRE steps and converged step count (24 atoms)
  terms.npz           the reference's sample_* term tables
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402
from firecore_amd import native  # noqa: E402

FIRE_REF = (0.05, 0.005, 0.1, 1.1, 0.5, 0.8, 3)


def pack(m):
    return {k: m[k] for k in ("apos", "REQs", "neighs", "apars", "bLs", "bKs", "Ksp", "Kpp", "elem")} | \
        {"natoms": np.int32(m["natoms"]), "nnode": np.int32(m["nnode"])}


def main():
    R = oracle.reference()
    assert R is not None, "build oracle/_ref first (make -C oracle ref)"
    specs = [(16, 1), (24, 0), (40, 2), (77, 3), (128, 4)]
    out = {}
    bk = {}
    for k, (n, s) in enumerate(specs):
        m = native.gen_molecule(n, s, 0.1, s + 1000)
        f, fng, fngpi, E = R.eval(m)
        for key, v in pack(m).items():
            out["m%d_%s" % (k, key)] = v
        out["m%d_f" % k], out["m%d_fng" % k], out["m%d_fngpi" % k], out["m%d_E" % k] = f, fng, fngpi, np.float64(E)
        r, b, ng = R.make_back_neighs(m["natoms"], m["nnode"], m["neighs"])
        bk["m%d_neighs" % k], bk["m%d_bk" % k], bk["m%d_neighs_out" % k] = m["neighs"], b, ng
        bk["m%d_nnode" % k] = np.int32(m["nnode"])
    np.savez_compressed(os.path.join(HERE, "forces_vacuum.npz"), n=len(specs), **out)
    np.savez_compressed(os.path.join(HERE, "backneighs.npz"), n=len(specs), **bk)

    # small grid (fixture size): 16 x 16 x 48 over the 5.64 A cell, dz = 0.2
    g = native.make_nacl_grid(16, 16, 48, 0.2, 8)
    gout = dict(gn=np.array(g["n"], np.int32), gpos0=g["pos0"], gstep=g["step"], galpha=np.float64(g["alpha"]),
                gcoefs=g["coefs"])
    rng = np.random.default_rng(7)
    P = np.stack([rng.uniform(-6, 6, 64), rng.uniform(-6, 6, 64), rng.uniform(0.2, 9.0, 64)], 1)
    PLQ = np.stack([rng.uniform(5, 40, 64), rng.uniform(0.1, 2, 64), rng.uniform(-0.5, 0.5, 64),
                    np.zeros(64)], 1)
    Es, Fs = [], []
    for p, q in zip(P, PLQ):
        E, f = R.grid_sample(g, p, q)
        Es.append(E)
        Fs.append(f)
    np.savez_compressed(os.path.join(HERE, "grid_samples.npz"), **gout, P=P, PLQ=PLQ, E=np.array(Es), F=np.array(Fs))
    gf = dict(gout)
    for k, (n, s) in enumerate([(30, 11), (60, 12)]):
        m = native.gen_molecule(n, s, 0.1, s + 1000)
        m["apos"][:n, 2] += 3.0 - m["apos"][:n, 2].min()
        f, fng, fngpi, E = R.eval(m, grid=g)
        for key, v in pack(m).items():
            gf["m%d_%s" % (k, key)] = v
        gf["m%d_f" % k], gf["m%d_E" % k] = f, np.float64(E)
    np.savez_compressed(os.path.join(HERE, "forces_grid.npz"), n=2, **gf)

    m = native.gen_molecule(24, 0, 0.1, 1)
    n0, apos0, E0, traj = R.relax(m, FIRE_REF, 1e-4, 5000, 0, ntraj=40)
    n1, apos1, E1, _ = R.relax(m, FIRE_REF, 1e-4, 5000, 1)
    np.savez_compressed(os.path.join(HERE, "fire_traj.npz"), **pack(m), traj=traj, nsteps_fmax=np.int32(n0),
                        E_fmax=np.float64(E0), apos_fmax=apos0, nsteps_ref=np.int32(n1), E_ref=np.float64(E1),
                        fire=np.array(FIRE_REF))

    ang = np.linspace(0.05, np.pi - 0.05, 25)
    terms = {"angles": ang}
    for kind in range(3):
        E, F = R.sample_terms(kind, 1.7, 2.0, 1.1, 1.3, ang)
        terms["E%d" % kind], terms["F%d" % kind] = E, F
    rs = np.linspace(1.5, 8.0, 40)
    REQij = np.array([3.4, 0.004, -0.03, 0.0])
    terms["rs"], terms["REQij"] = rs, REQij
    terms["Epair"], terms["Fpair"] = R.sample_pair(rs, REQij)
    np.savez_compressed(os.path.join(HERE, "terms.npz"), **terms)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
