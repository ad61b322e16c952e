"""Writes tests/golden/typed/ (inputs) and tests/golden/typed_build.npz
(expected outputs) for the typed-molecule builder behind init(xyz_name).

Expected outputs come from the reference's own MolWorld_sp3::init of each
molecule (MolWorld_sp3.h:309-357 with bMMFF, no electron pairs; run through
oracle/_ref/libfcref.so, built from /root/reference by oracle/Makefile), one
molecule per child process because the reference ends its process on a bad
input.  Run where the reference exists:

    make -C oracle all ref && python tests/golden/make_golden_builder.py

Inputs copied next to the outputs, as data: the reference's type tables
(common_resources/ElementTypes.dat, AtomTypes.dat, BondTypes.dat) and the
molecules' .xyz files (common_resources/xyz/).  Per molecule the npz holds
natoms, nnode, apos (atoms then pi-vectors), REQs, neighs, neighCell (image
of each bonded neighbour for init's nPBC = (1,1,0)), atypes, apars, bLs, bKs,
Ksp, Kpp and lvec (zeros for a molecule without an lvs line).
"""
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
RES = "/root/reference/cpp"
TABLES = ("ElementTypes.dat", "AtomTypes.dat", "BondTypes.dat")
# the molecules of the reference's MMFF_multi examples (tests/tMMFFmulti/run.py:
# PTCDA, pyridine, H2O, nHexadecan_dicarboxylic) and a spread of N, O, sp1/sp2/sp3
# chemistry; periodic ones (lvs line): PTCDA, pyridine, BPBA, formic_dimer,
# polydiacetylene, oxalate
MOLECULES = ("PTCDA", "pyridine", "H2O", "nHexadecan_dicarboxylic", "BPBA", "formic_dimer", "polydiacetylene",
             "oxalate", "guanine", "uracil", "CHONH2", "HCOOH_xy", "Benzene_deriv", "xylitol", "Si10_H",
             "N_2edge_assay_1", "CH2NH", "CO", "HCl", "NH3")
KEYS = ("apos", "REQs", "neighs", "neighCell", "atypes", "apars", "bLs", "bKs", "Ksp", "Kpp")


def one(name, out):
    sys.path.insert(0, ROOT)
    import oracle
    r, lv = oracle.reference().build_xyz_ref(RES, "common_resources/xyz/" + name, (1, 1, 0))
    np.savez(out, natoms=r["natoms"], nnode=r["nnode"], lvec=np.zeros((3, 3)) if lv is None else lv,
             **{k: r[k] for k in KEYS})


def main():
    dst = os.path.join(HERE, "typed")
    os.makedirs(dst, exist_ok=True)
    for t in TABLES:
        shutil.copyfile(os.path.join(RES, "common_resources", t), os.path.join(dst, t))
    z = {}
    tmp = os.path.join(dst, "_one.npz")
    for name in MOLECULES:
        shutil.copyfile(os.path.join(RES, "common_resources/xyz", name + ".xyz"), os.path.join(dst, name + ".xyz"))
        subprocess.run([sys.executable, __file__, "--one", name, tmp], check=True, stdout=subprocess.DEVNULL)
        with np.load(tmp) as d:
            for k in d.files:
                z[name + "/" + k] = d[k]
        os.remove(tmp)
        print(name, int(z[name + "/natoms"]), "atoms", int(z[name + "/nnode"]), "nodes")
    z["names"] = np.array(MOLECULES)
    np.savez_compressed(os.path.join(HERE, "typed_build.npz"), **z)


if __name__ == "__main__":
    if len(sys.argv) == 4 and sys.argv[1] == "--one":
        one(sys.argv[2], sys.argv[3])
    else:
        main()
