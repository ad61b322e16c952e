// Test-only driver around the reference's own CPU classes.
//
// This file is compiled by oracle/Makefile against the reference headers that
// live under /root/reference/cpp/common (nothing of the reference is copied
// here) into oracle/_ref/libfcref.so.  It calls the reference's code for
// every physics/optimizer step so that the FP64 oracle (oracle/oracle.cpp)
// can be pinned against it:
//   MMFFsp3_loc::eval_atom / assemble_atom / makeBackNeighs / move_atom_FIRE
//   NBFF::evalLJQs_ng4_atom_omp, GridFF::addAtom -> getForce_Bspline,
//   REQ2PLQ_d, DynamicOpt::FIRE_update_params
// The per-step sequence is the one of MolWorld_sp3::run_no_omp
// (MolWorld_sp3.h:2291-2474) with the FIRE branch.
#include "globals.h"
#include "testUtils.h"
#include "MMFFsp3_loc.h"
#include "GridFF.h"
#include "DynamicOpt.h"
#include "MolWorld_sp3.h"

#include <unistd.h>

#include <cmath>
#include <cstring>
#include <vector>

namespace {

// periodic molecule for the following ref_eval / ref_relax calls (ref_set_pbc)
struct RefPbc {
    bool on = false;
    double lvec[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int nPBC[3] = {0, 0, 0};
};
RefPbc g_ref_pbc;

// explicit substrate atoms for the following calls (ref_set_surface)
struct RefSurf {
    int natoms = 0;
    std::vector<Vec3d> apos;
    std::vector<Quat4d> REQs;
    std::vector<int> atypes;
    double lvec[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int nPBC[3] = {0, 0, 0};
    double alpha = 1.5;
};
RefSurf g_ref_surf;

struct RefSys {
    MMFFsp3_loc ffl;
    GridFF gff;
    bool bGrid = false;
    bool bNonBonded = true;
    std::vector<Quat4d> PLQd;
    DynamicOpt opt;

    void setup(int natoms, int nnode, const double* apos, const double* REQs, const int* neighs,
               const double* apars, const double* bLs, const double* bKs, const double* Ksp, const double* Kpp,
               const int* sw) {
        ffl.realloc(nnode, natoms - nnode);
        ffl.bPBC = false;
        for (int i = 0; i < ffl.nvecs; i++) ffl.apos[i] = Vec3d{apos[i * 3], apos[i * 3 + 1], apos[i * 3 + 2]};
        for (int i = 0; i < natoms; i++) {
            ffl.REQs[i] = Quat4d{REQs[i * 4], REQs[i * 4 + 1], REQs[i * 4 + 2], REQs[i * 4 + 3]};
            ffl.neighs[i] = Quat4i{neighs[i * 4], neighs[i * 4 + 1], neighs[i * 4 + 2], neighs[i * 4 + 3]};
            ffl.neighCell[i] = Quat4i{-1, -1, -1, -1};
        }
        for (int i = 0; i < nnode; i++) {
            ffl.apars[i] = Quat4d{apars[i * 4], apars[i * 4 + 1], apars[i * 4 + 2], apars[i * 4 + 3]};
            ffl.bLs[i] = Quat4d{bLs[i * 4], bLs[i * 4 + 1], bLs[i * 4 + 2], bLs[i * 4 + 3]};
            ffl.bKs[i] = Quat4d{bKs[i * 4], bKs[i * 4 + 1], bKs[i * 4 + 2], bKs[i * 4 + 3]};
            ffl.Ksp[i] = Quat4d{Ksp[i * 4], Ksp[i * 4 + 1], Ksp[i * 4 + 2], Ksp[i * 4 + 3]};
            ffl.Kpp[i] = Quat4d{Kpp[i * 4], Kpp[i * 4 + 1], Kpp[i * 4 + 2], Kpp[i * 4 + 3]};
        }
        ffl.doBonds = sw[0];
        ffl.doPiPiI = sw[1];
        ffl.doPiSigma = sw[2];
        ffl.doAngles = sw[3];
        bNonBonded = sw[4];
        ffl.bNonBonded = bNonBonded;
        ffl.setNonBondStrategy(+1);  // bNonBondNeighs
        ffl.makeBackNeighs();
        if (g_ref_pbc.on) {  // MolWorld_sp3::makeMMFFs PBC set-up (MolWorld_sp3.h:1416-1432)
            // npbc = makePBCshifts(nPBC, lvec); ffl.bindShifts(npbc, shifts);
            // ffl.makeNeighCells(npbc, shifts) (NBFF::makePBCshifts returns the
            // count without storing it, so it is bound as MolWorld_sp3 does)
            ffl.setLvec(*(const Mat3d*)g_ref_pbc.lvec);
            const int npbc = ffl.makePBCshifts(Vec3i{g_ref_pbc.nPBC[0], g_ref_pbc.nPBC[1], g_ref_pbc.nPBC[2]}, true);
            ffl.bindShifts(npbc, ffl.shifts);
            ffl.makeNeighCells(npbc, ffl.shifts);
            ffl.bPBC = true;
        }
    }
    void setup_grid(int nx, int ny, int nz, const double* pos0, const double* step, double* coefs, double alpha) {
        bGrid = true;
        gff.grid.n = Vec3i{nx, ny, nz};
        gff.grid.pos0 = Vec3d{pos0[0], pos0[1], pos0[2]};
        gff.grid.diCell = Mat3d{1.0 / step[0], 0, 0, 0, 1.0 / step[1], 0, 0, 0, 1.0 / step[2]};
        gff.Bspline_PLQ = (Vec3d*)coefs;
        gff.mode = GridFFmod::BsplineDouble;
        make_inds_pbc(nx, gff.cubic_xqis);
        make_inds_pbc(ny, gff.cubic_yqis);
        gff.alphaMorse = alpha;
        PLQd.resize(ffl.natoms);
        for (int i = 0; i < ffl.natoms; i++) PLQd[i] = REQ2PLQ_d(ffl.REQs[i], alpha);
    }
    // substrate atoms: MolWorld_sp3::loadSurf without a grid (MolWorld_sp3.h:
    // 1072-1075): gridFF.bindSystem, gridFF.lvec, makePBCshifts(nPBC, lvec)
    GridFF gsurf;
    bool bSurf = false;
    void setup_surf() {
        if (g_ref_surf.natoms <= 0) return;
        bSurf = true;
        gsurf.bindSystem(g_ref_surf.natoms, g_ref_surf.atypes.data(), g_ref_surf.apos.data(), g_ref_surf.REQs.data());
        gsurf.lvec = *(const Mat3d*)g_ref_surf.lvec;
        gsurf.makePBCshifts(Vec3i{g_ref_surf.nPBC[0], g_ref_surf.nPBC[1], g_ref_surf.nPBC[2]}, gsurf.lvec);
        gsurf.alphaMorse = g_ref_surf.alpha;
        gsurf.shift0 = Vec3dZero;
    }
    ~RefSys() {  // the substrate arrays belong to g_ref_surf
        gsurf.natoms = 0;
        gsurf.apos = nullptr;
        gsurf.REQs = nullptr;
        gsurf.atypes = nullptr;
    }
    // MolWorld_sp3::eval_no_omp (MolWorld_sp3.h:2178-2245) recipe
    double eval() {
        double E = 0;
        for (int i = 0; i < ffl.natoms; i++) ffl.fapos[i] = Vec3dZero;
        for (int ia = 0; ia < ffl.natoms; ia++) {
            ffl.fapos[ia] = Vec3dZero;
            if (ia < ffl.nnode) ffl.fapos[ia + ffl.natoms] = Vec3dZero;
            if (ia < ffl.nnode) E += ffl.eval_atom(ia);
            if (bNonBonded) E += ffl.bPBC ? ffl.evalLJQs_ng4_PBC_atom_omp(ia) : ffl.evalLJQs_ng4_atom_omp(ia);
            if (bGrid) E += gff.addAtom(ffl.apos[ia], PLQd[ia], ffl.fapos[ia]);
            else if (bSurf) E += gsurf.evalMorsePBC_sym(ffl.apos[ia], ffl.REQs[ia], ffl.fapos[ia]);
        }
        for (int ia = 0; ia < ffl.natoms; ia++) ffl.assemble_atom(ia);
        return E;
    }
};

void copy_out(const RefSys& R, double* fapos, double* fneigh, double* fneighpi) {
    if (fapos)
        for (int i = 0; i < R.ffl.nvecs; i++) {
            fapos[i * 3] = R.ffl.fapos[i].x;
            fapos[i * 3 + 1] = R.ffl.fapos[i].y;
            fapos[i * 3 + 2] = R.ffl.fapos[i].z;
        }
    if (fneigh)
        for (int i = 0; i < R.ffl.nnode * 4; i++) {
            fneigh[i * 3] = R.ffl.fneigh[i].x;
            fneigh[i * 3 + 1] = R.ffl.fneigh[i].y;
            fneigh[i * 3 + 2] = R.ffl.fneigh[i].z;
        }
    if (fneighpi)
        for (int i = 0; i < R.ffl.nnode * 4; i++) {
            fneighpi[i * 3] = R.ffl.fneighpi[i].x;
            fneighpi[i * 3 + 1] = R.ffl.fneighpi[i].y;
            fneighpi[i * 3 + 2] = R.ffl.fneighpi[i].z;
        }
}

}  // namespace

extern "C" {

// Explicit substrate atoms (positions, REQs, lattice rows, image range, Morse
// alpha) for the following calls; natoms = 0 removes them.
void ref_set_surface(int natoms, const double* apos, const double* REQs, const double* lvec9, const int* nPBC3,
                     double alpha) {
    g_ref_surf.natoms = natoms;
    if (natoms <= 0) return;
    g_ref_surf.apos.resize(natoms);
    g_ref_surf.REQs.resize(natoms);
    g_ref_surf.atypes.assign(natoms, 0);
    for (int i = 0; i < natoms; i++) {
        g_ref_surf.apos[i] = Vec3d{apos[i * 3], apos[i * 3 + 1], apos[i * 3 + 2]};
        g_ref_surf.REQs[i] = Quat4d{REQs[i * 4], REQs[i * 4 + 1], REQs[i * 4 + 2], REQs[i * 4 + 3]};
    }
    for (int k = 0; k < 9; k++) g_ref_surf.lvec[k] = lvec9[k];
    for (int k = 0; k < 3; k++) g_ref_surf.nPBC[k] = nPBC3[k];
    g_ref_surf.alpha = alpha;
}

// Periodic molecule (lattice rows, image range) for the following calls;
// lvec9 = null switches back to vacuum molecules.
void ref_set_pbc(const double* lvec9, const int* nPBC3) {
    g_ref_pbc.on = lvec9 != nullptr;
    if (!lvec9) return;
    for (int k = 0; k < 9; k++) g_ref_pbc.lvec[k] = lvec9[k];
    for (int k = 0; k < 3; k++) g_ref_pbc.nPBC[k] = nPBC3[k];
}

// The reference's own MMFF build of an .xyz molecule: MolWorld_sp3::init with
// xyz_name and bMMFF (MolWorld_sp3.h:309-357), run with the working directory
// set to res_dir (where its common_resources/ type tables live).  Copies out
// the MMFFsp3_loc arrays: apos (nvecs x 3), REQs (natoms x 4), neighs and
// neighCell (natoms x 4), apars/bLs/bKs/Ksp/Kpp (nnode x 4), atom types
// (natoms), lattice (9) and the periodic flag.  Returns natoms (nnode in
// *nnode_out), -1 if it does not fit maxn.  The world is left allocated: its
// destructor is not run.
int ref_build_xyz(const char* res_dir, const char* xyz_noext, const int* nPBC3, int maxn, int* nnode_out,
                  double* apos, double* REQs, int* neighs, int* neighCell, double* apars, double* bLs, double* bKs,
                  double* Ksp, double* Kpp, int* atypes, double* lvec9, int* periodic) {
    char cwd[4096];
    if (!getcwd(cwd, sizeof(cwd))) return -1;
    if (chdir(res_dir) != 0) return -1;
    MolWorld_sp3* W = new MolWorld_sp3();
    W->xyz_name = xyz_noext;
    W->bMMFF = true;
    W->bEpairs = false;
    W->nPBC = Vec3i{nPBC3[0], nPBC3[1], nPBC3[2]};
    W->init();
    if (chdir(cwd) != 0) return -1;
    const MMFFsp3_loc& f = W->ffl;
    const int N = f.natoms, M = f.nnode;
    if (N > maxn) return -1;
    *nnode_out = M;
    for (int i = 0; i < f.nvecs; i++) {
        apos[i * 3] = f.apos[i].x;
        apos[i * 3 + 1] = f.apos[i].y;
        apos[i * 3 + 2] = f.apos[i].z;
    }
    for (int i = 0; i < N; i++) {
        for (int k = 0; k < 4; k++) {
            REQs[i * 4 + k] = f.REQs[i].array[k];
            neighs[i * 4 + k] = f.neighs[i].array[k];
            neighCell[i * 4 + k] = f.neighCell[i].array[k];
        }
        atypes[i] = f.atypes[i];
    }
    for (int i = 0; i < M; i++)
        for (int k = 0; k < 4; k++) {
            apars[i * 4 + k] = f.apars[i].array[k];
            bLs[i * 4 + k] = f.bLs[i].array[k];
            bKs[i * 4 + k] = f.bKs[i].array[k];
            Ksp[i * 4 + k] = f.Ksp[i].array[k];
            Kpp[i * 4 + k] = f.Kpp[i].array[k];
        }
    const Mat3d& L = W->builder.lvec;
    const double lv[9] = {L.a.x, L.a.y, L.a.z, L.b.x, L.b.y, L.b.z, L.c.x, L.c.y, L.c.z};
    for (int k = 0; k < 9; k++) lvec9[k] = lv[k];
    *periodic = W->bPBC ? 1 : 0;
    return N;
}

// Image shifts (NBFF::makePBCshifts) and neighbour images
// (MMFFsp3_loc::makeNeighCells(npbc, shifts)) of a periodic molecule.
int ref_neigh_cells(int natoms, int nnode, const double* apos, const int* neighs, const double* lvec9,
                    const int* nPBC3, int* neighCell_out, double* shifts_out) {
    MMFFsp3_loc ffl;
    ffl.realloc(nnode, natoms - nnode);
    for (int i = 0; i < ffl.nvecs; i++) ffl.apos[i] = Vec3d{apos[i * 3], apos[i * 3 + 1], apos[i * 3 + 2]};
    for (int i = 0; i < natoms; i++)
        ffl.neighs[i] = Quat4i{neighs[i * 4], neighs[i * 4 + 1], neighs[i * 4 + 2], neighs[i * 4 + 3]};
    ffl.setLvec(*(const Mat3d*)lvec9);
    const int npbc = ffl.makePBCshifts(Vec3i{nPBC3[0], nPBC3[1], nPBC3[2]}, true);
    ffl.makeNeighCells(npbc, ffl.shifts);
    for (int i = 0; i < natoms; i++)
        for (int k = 0; k < 4; k++) neighCell_out[i * 4 + k] = ffl.neighCell[i].array[k];
    for (int i = 0; i < npbc; i++) {
        shifts_out[i * 3] = ffl.shifts[i].x;
        shifts_out[i * 3 + 1] = ffl.shifts[i].y;
        shifts_out[i * 3 + 2] = ffl.shifts[i].z;
    }
    return npbc;
}

int ref_make_back_neighs(int natoms, int nnode, const int* neighs, int* bk_out, int* neighs_out) {
    MMFFsp3_loc ffl;
    ffl.realloc(nnode, natoms - nnode);
    for (int i = 0; i < natoms; i++)
        ffl.neighs[i] = Quat4i{neighs[i * 4], neighs[i * 4 + 1], neighs[i * 4 + 2], neighs[i * 4 + 3]};
    ffl.makeBackNeighs();
    for (int i = 0; i < natoms; i++) {
        for (int k = 0; k < 4; k++) {
            bk_out[i * 4 + k] = ffl.bkneighs[i].array[k];
            if (neighs_out) neighs_out[i * 4 + k] = ffl.neighs[i].array[k];
        }
    }
    return 0;
}

// Forces (assembled, atoms then pi), recoils and total energy of one system.
double ref_eval(int natoms, int nnode, const double* apos, const double* REQs, const int* neighs, const double* apars,
                const double* bLs, const double* bKs, const double* Ksp, const double* Kpp, const int* sw, int gnx,
                int gny, int gnz, const double* gpos0, const double* gstep, double* gcoefs, double alpha,
                double* fapos, double* fneigh, double* fneighpi) {
    RefSys R;
    R.setup(natoms, nnode, apos, REQs, neighs, apars, bLs, bKs, Ksp, Kpp, sw);
    R.setup_surf();
    if (gcoefs) R.setup_grid(gnx, gny, gnz, gpos0, gstep, gcoefs, alpha);
    double E = R.eval();
    copy_out(R, fapos, fneigh, fneighpi);
    if (gcoefs) R.gff.Bspline_PLQ = nullptr;
    return E;
}

// B-spline GridFF sample of one point with coefficients PLQ (GridFF::addAtom).
double ref_grid_sample(int gnx, int gny, int gnz, const double* gpos0, const double* gstep, double* gcoefs,
                       const double* p, const double* plq, double* f) {
    GridFF gff;
    gff.grid.n = Vec3i{gnx, gny, gnz};
    gff.grid.pos0 = Vec3d{gpos0[0], gpos0[1], gpos0[2]};
    gff.grid.diCell = Mat3d{1.0 / gstep[0], 0, 0, 0, 1.0 / gstep[1], 0, 0, 0, 1.0 / gstep[2]};
    gff.Bspline_PLQ = (Vec3d*)gcoefs;
    gff.mode = GridFFmod::BsplineDouble;
    make_inds_pbc(gnx, gff.cubic_xqis);
    make_inds_pbc(gny, gff.cubic_yqis);
    Vec3d fo = Vec3dZero;
    double E = gff.addAtom(Vec3d{p[0], p[1], p[2]}, Quat4d{plq[0], plq[1], plq[2], plq[3]}, fo);
    f[0] = fo.x;
    f[1] = fo.y;
    f[2] = fo.z;
    gff.Bspline_PLQ = nullptr;
    return E;
}

// FIRE relaxation of one system with the run_no_omp recipe.
//   fire = {dt_max, dt_min, damp_max, finc, fdec, falpha, minLastNeg}
//   conv_mode 0: stop before moving when max_i |f_i| < Fconv (all vecs)
//   conv_mode 1: reference criterion, sum f^2 < Fconv^2 tested after the move
// apos is updated in place; traj (optional) receives the positions after each
// of the first ntraj steps.  Returns the step index at convergence (nmax if
// not converged); E_out receives the energy of the last evaluation.
int ref_relax(int natoms, int nnode, double* apos, const double* REQs, const int* neighs, const double* apars,
              const double* bLs, const double* bKs, const double* Ksp, const double* Kpp, const int* sw, int gnx,
              int gny, int gnz, const double* gpos0, const double* gstep, double* gcoefs, double alpha,
              const double* fire, double Fconv, int nmax, int conv_mode, double* E_out, double* traj, int ntraj) {
    RefSys R;
    R.setup(natoms, nnode, apos, REQs, neighs, apars, bLs, bKs, Ksp, Kpp, sw);
    R.setup_surf();
    if (gcoefs) R.setup_grid(gnx, gny, gnz, gpos0, gstep, gcoefs, alpha);
    MMFFsp3_loc& ffl = R.ffl;
    DynamicOpt& opt = R.opt;
    opt.bindOrAlloc(ffl.nDOFs, ffl.DOFs, 0, ffl.fDOFs, 0);
    opt.cleanVel();
    ffl.vapos = (Vec3d*)opt.vel;
    opt.dt = opt.dt_max = fire[0];
    opt.dt_min = fire[1];
    opt.damping = opt.damp_max = fire[2];
    opt.finc = fire[3];
    opt.fdec = fire[4];
    opt.falpha = fire[5];
    opt.minLastNeg = (int)fire[6];
    opt.lastNeg = 0;
    const double F2conv = Fconv * Fconv;
    int itr = 0;
    double E = 0;
    for (itr = 0; itr < nmax; itr++) {
        E = R.eval();
        double vf = 0, vv = 0, ff = 0, fm2 = 0;
        for (int i = 0; i < opt.n; i++) {
            double v = opt.vel[i], f = opt.force[i];
            vf += v * f;
            vv += v * v;
            ff += f * f;
        }
        for (int i = 0; i < ffl.nvecs; i++) fm2 = std::max(fm2, ffl.fapos[i].norm2());
        if (conv_mode == 0 && fm2 < F2conv) break;
        opt.vf = vf;
        opt.vv = vv;
        opt.ff = ff;
        opt.FIRE_update_params();
        for (int i = 0; i < ffl.nvecs; i++) ffl.move_atom_FIRE(i, opt.dt, 10000.0, opt.cv, opt.renorm_vf * opt.cf);
        if (traj && itr < ntraj)
            for (int i = 0; i < ffl.nvecs; i++) {
                traj[((size_t)itr * ffl.nvecs + i) * 3] = ffl.apos[i].x;
                traj[((size_t)itr * ffl.nvecs + i) * 3 + 1] = ffl.apos[i].y;
                traj[((size_t)itr * ffl.nvecs + i) * 3 + 2] = ffl.apos[i].z;
            }
        if (conv_mode == 1 && ff < F2conv) break;
    }
    for (int i = 0; i < ffl.nvecs; i++) {
        apos[i * 3] = ffl.apos[i].x;
        apos[i * 3 + 1] = ffl.apos[i].y;
        apos[i * 3 + 2] = ffl.apos[i].z;
    }
    if (E_out) *E_out = E;
    if (gcoefs) R.gff.Bspline_PLQ = nullptr;
    return itr;
}

// The reference's own sample_* term harness semantics (libMMFF.h:439-473):
// h1=(1,0,0), h2=(cos a, sin a, 0); returns E and f1.y.
void ref_sample_terms(int kind, double k, double ang0, double r1, double r2, int n, const double* angles, double* Es,
                      double* Fs) {
    Vec3d h1{1, 0, 0}, f1, f2;
    Vec2d cs0;
    cs0.fromAngle(ang0 / 2.);
    for (int i = 0; i < n; i++) {
        Vec3d h2{cos(angles[i]), sin(angles[i]), 0.0};
        if (kind == 0) Es[i] = evalPiAling(h1, h2, 1. / r1, 1. / r2, k, f1, f2);
        else if (kind == 1) Es[i] = evalAngleCos(h1, h2, 1. / r1, 1. / r2, k, cos(ang0), f1, f2);
        else Es[i] = evalAngleCosHalf(h1, h2, 1. / r1, 1. / r2, cs0, k, f1, f2);
        Fs[i] = f1.y;
    }
}

// getLJQH pair energy/force along x (Forces.h:341-360); F = force on the +x atom.
void ref_sample_pair(int n, const double* rs, const double* REQij, double* Es, double* Fs) {
    Quat4d q{REQij[0], REQij[1], REQij[2], REQij[3]};
    for (int i = 0; i < n; i++) {
        Vec3d f;
        Es[i] = getLJQH(Vec3d{rs[i], 0, 0}, f, q, 1.0);
        Fs[i] = -f.x;
    }
}

// MMFFsp3_loc::initPi (MMFFsp3_loc.h:882-945) on a non-periodic molecule
// (every neighbour in cell 0 with a zero shift), optionally followed by
// flipPis(Vec3dOne) (MMFFsp3_loc.h:1167-1172) as MolWorld_sp3 does after
// building (MolWorld_sp3.h:1400).  pipos_out: nnode x 3.
void ref_init_pi(int natoms, int nnode, const double* apos, const int* neighs, const double* Ksp, const double* Kpp,
                 int flip, double* pipos_out) {
    MMFFsp3_loc ffl;
    ffl.realloc(nnode, natoms - nnode);
    ffl.bPBC = false;
    for (int i = 0; i < ffl.nvecs; i++) ffl.apos[i] = Vec3dZero;
    for (int i = 0; i < natoms; i++) {
        ffl.apos[i] = Vec3d{apos[i * 3], apos[i * 3 + 1], apos[i * 3 + 2]};
        ffl.neighs[i] = Quat4i{neighs[i * 4], neighs[i * 4 + 1], neighs[i * 4 + 2], neighs[i * 4 + 3]};
        ffl.neighCell[i] = Quat4i{0, 0, 0, 0};
    }
    for (int i = 0; i < nnode; i++) {
        ffl.Ksp[i] = Quat4d{Ksp[i * 4], Ksp[i * 4 + 1], Ksp[i * 4 + 2], Ksp[i * 4 + 3]};
        ffl.Kpp[i] = Quat4d{Kpp[i * 4], Kpp[i * 4 + 1], Kpp[i * 4 + 2], Kpp[i * 4 + 3]};
    }
    Vec3d shift0 = Vec3dZero;
    ffl.initPi(&shift0);
    if (flip) ffl.flipPis(Vec3dOne);
    for (int i = 0; i < nnode; i++) {
        pipos_out[i * 3] = ffl.pipos[i].x;
        pipos_out[i * 3 + 1] = ffl.pipos[i].y;
        pipos_out[i * 3 + 2] = ffl.pipos[i].z;
    }
}

// GridFF node potentials of a substrate at npts points with the reference's
// own GridFF::evalGridFFPoint_Mors (GridFF.h:582-611) over the images of
// makePBCshifts_({nmx,nmy,0}) and evalGridFFPoint_Coul (:614-635, bare) over
// those of ({ncx,ncy,0}) (NBFF.h:42-57): out[i] = (Pauli, London, Coulomb).
void ref_gridff_points(int natoms, const double* apos, const double* REQs, const double* lvec9, double alpha,
                       int nmx, int nmy, int ncx, int ncy, int npts, const double* pts, double* out) {
    GridFF gff;
    gff.alphaMorse = alpha;
    Mat3d lvec{lvec9[0], lvec9[1], lvec9[2], lvec9[3], lvec9[4], lvec9[5], lvec9[6], lvec9[7], lvec9[8]};
    Vec3d* sm = nullptr;
    Vec3d* sc = nullptr;
    const int nm = makePBCshifts_(Vec3i{nmx, nmy, 0}, lvec, sm);
    const int nc = makePBCshifts_(Vec3i{ncx, ncy, 0}, lvec, sc);
    std::vector<Vec3d> ap(natoms);
    std::vector<Quat4d> rq(natoms);
    for (int i = 0; i < natoms; i++) {
        ap[i] = Vec3d{apos[i * 3], apos[i * 3 + 1], apos[i * 3 + 2]};
        rq[i] = Quat4d{REQs[i * 4], REQs[i * 4 + 1], REQs[i * 4 + 2], REQs[i * 4 + 3]};
    }
    for (int i = 0; i < npts; i++) {
        const Vec3d p{pts[i * 3], pts[i * 3 + 1], pts[i * 3 + 2]};
        Quat4d qp, ql;
        gff.evalGridFFPoint_Mors(nm, sm, natoms, ap.data(), rq.data(), p, qp, ql);
        const Quat4d qe = gff.evalGridFFPoint_Coul(nc, sc, natoms, ap.data(), rq.data(), p);
        out[i * 3] = qp.w;
        out[i * 3 + 1] = ql.w;
        out[i * 3 + 2] = qe.w;
    }
    delete[] sm;
    delete[] sc;
}

// ---- the reference's MolecularDatabase (MolecularDatabase.h) with its default
// descriptors (principal axes of all atoms, number of atoms).  Structures are
// kept alive for the database's lifetime (it stores copies, but the candidate's
// arrays are read during addIfNewDescriptor).
static MolecularDatabase* g_db = nullptr;

// the arrays a structure was offered with, kept for the process lifetime
static void db_bind(ForceField& ff, int n, const double* pos, const int* types) {
    Vec3d* p = new Vec3d[n];
    int* t = new int[n];
    for (int i = 0; i < n; i++) {
        p[i] = Vec3d{pos[i * 3], pos[i * 3 + 1], pos[i * 3 + 2]};
        t[i] = types[i];
    }
    ff.natoms = n;
    ff.apos = p;
    ff.atypes = t;
    ff.vapos = new Vec3d[n]();
    ff.fapos = new Vec3d[n]();
}

void ref_db_reset() {
    g_db = new MolecularDatabase();  // the previous one is left allocated (test process only)
    Atoms dummy(3);
    for (int i = 0; i < 3; i++) {
        dummy.atypes[i] = 0;
        dummy.apos[i] = Vec3d{(double)i, 0, 0};
    }
    g_db->setDescriptors(&dummy);  // default descriptor set; reads atypes[0..2] of its argument
}

// addIfNewDescriptor: -1 = added as a new member, else the index it duplicates
int ref_db_add(int n, const double* pos, const int* types) {
    if (!g_db) ref_db_reset();
    ForceField ff;
    db_bind(ff, n, pos, types);
    const int r = g_db->addIfNewDescriptor(&ff);
    ff.apos = nullptr;
    ff.atypes = nullptr;
    ff.vapos = nullptr;
    ff.fapos = nullptr;
    return r;
}

int ref_db_count() { return g_db ? g_db->getNMembers() : 0; }

int ref_db_descriptor(int n, const double* pos, double* out) {
    std::vector<int> t(n, 0);
    Atoms a;
    std::vector<Vec3d> p(n);
    for (int i = 0; i < n; i++) p[i] = Vec3d{pos[i * 3], pos[i * 3 + 1], pos[i * 3 + 2]};
    a.bind(n, t.data(), p.data());
    Descriptor d;
    d.setDescriptor(MolecularDatabaseDescriptor{principal_axes, -1}, &a);
    d.setDescriptor(MolecularDatabaseDescriptor{number_of_atoms, -1}, &a);
    const int nd = d.dimensionOfDescriptorSpace;
    for (int k = 0; k < nd; k++) out[k] = d.descs[k];
    a.bind(0, nullptr, nullptr);
    return nd;
}

// computeDistance(x, y) of the database: positions of both are restored after
double ref_db_distance(int n, const double* x, const int* tx, const double* y, const int* ty) {
    if (!g_db) ref_db_reset();
    std::vector<Vec3d> px(n), py(n);
    std::vector<int> vx(tx, tx + n), vy(ty, ty + n);
    for (int i = 0; i < n; i++) {
        px[i] = Vec3d{x[i * 3], x[i * 3 + 1], x[i * 3 + 2]};
        py[i] = Vec3d{y[i * 3], y[i * 3 + 1], y[i * 3 + 2]};
    }
    Atoms ax, ay;
    ax.bind(n, vx.data(), px.data());
    ay.bind(n, vy.data(), py.data());
    const double d = g_db->computeDistance(&ax, &ay);
    ax.bind(0, nullptr, nullptr);
    ay.bind(0, nullptr, nullptr);
    return d;
}

}  // extern "C"
