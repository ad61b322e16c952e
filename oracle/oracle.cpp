// FP64 CPU oracle of the multi-system MMFF relaxation path.
//
// TEST INFRASTRUCTURE ONLY.  This is a plain restatement, in double
// precision, of the reference CPU path that the GPU library must reproduce.
// Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
// may load it, and only to check results or as the timed CPU baseline; the
// library never calls it.
//
// What each function restates (reference = ProkopHapala/FireCore):
//   orc_eval          MolWorld_sp3::eval_no_omp        MolWorld_sp3.h:2178-2245
//     eval_node       MMFFsp3_loc::eval_atom           MMFFsp3_loc.h:250-575
//                     (bNonBondNeighs strategy: 1-3 pairs subtracted in the
//                     angle loop, no 1-2 subtraction, no clamp,
//                     ForceField.h:168-184)
//     nb_atom         NBFF::evalLJQs_ng4_atom_omp      NBFF.h:666-691
//     pair_ljqh       getLJQH                          Forces.h:341-360
//     grid_sample     GridFF::getForce_Bspline         GridFF.h:355-402
//                     Bspline::fe3d_pbc_comb3          Bspline.h:742-809
//                     basis/dbasis                     Bspline.h:253-276
//                     make_inds_pbc/choose_inds_pbc    Bspline.h:39-90
//     REQ -> PLQ      REQ2PLQ_d                        Forces.h:611-619
//     assemble        MMFFsp3_loc::assemble_atom       MMFFsp3_loc.h:998-1021
//   orc_make_back_neighs  MMFFsp3_loc::makeBackNeighs  MMFFsp3_loc.h:1335-1358
//   orc_relax         MolWorld_sp3::run_no_omp (FIRE) MolWorld_sp3.h:2261-2474
//     fire_update     DynamicOpt::FIRE_update_params   DynamicOpt.cpp:250-267
//     move            MMFFsp3_loc::move_atom_FIRE      MMFFsp3_loc.h:1264-1281
//                     (dt passed as float, renorm applied twice as in the
//                     reference call at MolWorld_sp3.h:2426)
//   orc_relax_batch   one system per thread, dynamic schedule, the structure
//                     of MolWorld_sp3_multi::run_omp_ocl (:1948-2027)
//   orc_relax_batch_f32  the same in single precision (the FP32 CPU baseline;
//                     not a parity reference)
//
// The restatement itself is oracle_core.h, written for a scalar `real` and
// included twice: namespace orc64 (double: every entry point below except
// orc_relax_batch_f32) and orc32 (float).
//   orc_set_surface   explicit substrate atoms: surf_atom restates
//                     GridFF::evalMorsePBC (GridFF.h:1229-1261)
//   orc_set_pbc       periodic molecules: nb_atom_pbc restates
//                     NBFF::evalLJQs_ng4_PBC_atom_omp (NBFF.h:575-620), bonds
//                     at their neighCell image (MMFFsp3_loc.h:325-329)
//   orc_explore       MolWorld_sp3_multi::updateMultiExploring (phase
//                     machine, MolWorld_sp3_multi.h:770-833) with the TDrive
//                     Langevin term of updateAtomsMMFFf4 (relax_multi.cl:
//                     1160-1178); random numbers from a counter-based hash
//
// Build: oracle/Makefile -> oracle/liboracle.so
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include <xmmintrin.h>

// the core twice: FP64 (the oracle) and FP32 (the FP32 CPU baseline)
namespace orc64 {
using real = double;
#include "oracle_core.h"
}  // namespace orc64
namespace orc32 {
using real = float;
#include "oracle_core.h"
}  // namespace orc32
using namespace orc64;

extern "C" {

// Exploring run of one system (explore_sys); apos updated in place.  Returns
// the number of minima found, -1 if the system blew up.
int orc_explore(int natoms, int nnode, double* apos, const double* REQs, const int* neighs, const double* apars,
                const double* bLs, const double* bKs, const double* Ksp, const double* Kpp, const int* sw, int gnx,
                int gny, int gnz, const double* gpos0, const double* gstep, const double* gcoefs, double alpha,
                const double* fire, double Fconv, int nmax, double T, double gamma, int nExplore, int nRelax,
                unsigned seed, int sysidx, int kmax, double* minE, double* minpos, double* traj, int ntraj) {
    Sys S;
    setup(S, natoms, nnode, apos, REQs, neighs, apars, bLs, bKs, Ksp, Kpp, sw, gnx, gny, gnz, gpos0, gstep, gcoefs,
          alpha);
    int n = explore_sys(S, fire, Fconv, nmax, T, gamma, nExplore, nRelax, seed, (uint32_t)sysidx, kmax, minE, minpos,
                        traj, ntraj);
    std::memcpy(apos, S.pos.data(), sizeof(double) * S.pos.size());
    return n;
}

// The exploring drive's random numbers, for the tests of its statistics.
double orc_drive_xi(unsigned seed, unsigned sys, unsigned vec, unsigned axis, unsigned step) {
    return drive_xi(seed, sys, vec, axis, step);
}

int orc_make_back_neighs(int natoms, int nnode, const int* neighs, int* bk, int* neighs_out) {
    for (int i = 0; i < natoms * 4; i++) bk[i] = -1;
    for (int ia = 0; ia < nnode; ia++)
        for (int j = 0; j < 4; j++) {
            int ja = neighs[ia * 4 + j];
            if (ja < 0) continue;
            int* row = bk + ja * 4;
            int k = 0;
            while (k < 4 && row[k] >= 0) k++;
            if (k == 4) return -1;
            row[k] = ia * 4 + j;
        }
    if (neighs_out) {
        std::memcpy(neighs_out, neighs, sizeof(int) * natoms * 4);
        for (int ia = nnode; ia < natoms; ia++) {  // cap neighbours from back-neighbours (:1348-1355)
            if (bk[ia * 4] < 0) return -2;
            neighs_out[ia * 4] = bk[ia * 4] / 4;
            neighs_out[ia * 4 + 1] = neighs_out[ia * 4 + 2] = neighs_out[ia * 4 + 3] = -1;
        }
    }
    return 0;
}

void orc_set_surface(int natoms, const double* apos, const double* REQs, int npbc, const double* shifts,
                     double alpha) {
    g_surf.natoms = natoms;
    if (natoms <= 0) return;
    g_surf.npbc = npbc;
    g_surf.apos.assign(apos, apos + natoms * 3);
    g_surf.REQs.assign(REQs, REQs + natoms * 4);
    g_surf.shifts.assign(shifts, shifts + npbc * 3);
    g_surf.K = -alpha;
}

void orc_set_pbc(int natoms, int npbc, const double* shifts, const int* neighCell) {
    g_pbc.natoms = natoms;
    g_pbc.npbc = npbc;
    if (natoms <= 0 || npbc <= 0) return;
    g_pbc.shifts.assign(shifts, shifts + npbc * 3);
    g_pbc.ngC.assign(neighCell, neighCell + natoms * 4);
}

void orc_set_restraints(int natoms, const double* constr, const double* constrK, const double* bbox9) {
    g_restraints.natoms = natoms;
    if (natoms <= 0) return;
    g_restraints.constr.assign(constr, constr + natoms * 4);
    g_restraints.constrK.assign(constrK, constrK + natoms * 4);
    for (int k = 0; k < 9; k++) g_restraints.bbox[k] = bbox9 ? bbox9[k] : 0.0;
}

double orc_eval(int natoms, int nnode, const double* apos, const double* REQs, const int* neighs, const double* apars,
                const double* bLs, const double* bKs, const double* Ksp, const double* Kpp, const int* sw, int gnx,
                int gny, int gnz, const double* gpos0, const double* gstep, const double* gcoefs, double alpha,
                double* fapos, double* fneigh, double* fneighpi) {
    Sys S;
    setup(S, natoms, nnode, apos, REQs, neighs, apars, bLs, bKs, Ksp, Kpp, sw, gnx, gny, gnz, gpos0, gstep, gcoefs,
          alpha);
    double E = eval_sys(S);
    if (fapos) std::memcpy(fapos, S.f.data(), sizeof(double) * S.f.size());
    if (fneigh) std::memcpy(fneigh, S.fng.data(), sizeof(double) * S.fng.size());
    if (fneighpi) std::memcpy(fneighpi, S.fngpi.data(), sizeof(double) * S.fngpi.size());
    return E;
}

double orc_grid_sample(int gnx, int gny, int gnz, const double* gpos0, const double* gstep, const double* gcoefs,
                       const double* p, const double* plq, double* f) {
    Sys S{};
    S.gc = gcoefs;
    S.gn[0] = gnx;
    S.gn[1] = gny;
    S.gn[2] = gnz;
    for (int k = 0; k < 3; k++) {
        S.gp0[k] = gpos0[k];
        S.gstep[k] = gstep[k];
    }
    V fo{0, 0, 0};
    double E = grid_sample(S, ld(p), plq, fo);
    st(f, fo);
    return E;
}

void orc_req2plq(const double* REQ, double alpha, double* PLQ) { req2plq(REQ, alpha, PLQ); }

int orc_relax(int natoms, int nnode, double* apos, const double* REQs, const int* neighs, const double* apars,
              const double* bLs, const double* bKs, const double* Ksp, const double* Kpp, const int* sw, int gnx,
              int gny, int gnz, const double* gpos0, const double* gstep, const double* gcoefs, double alpha,
              const double* fire, double Fconv, int nmax, int conv_mode, double* E_out, double* traj, int ntraj) {
    Sys S;
    setup(S, natoms, nnode, apos, REQs, neighs, apars, bLs, bKs, Ksp, Kpp, sw, gnx, gny, gnz, gpos0, gstep, gcoefs,
          alpha);
    int n = relax_sys(S, fire, Fconv, nmax, conv_mode, E_out, traj, ntraj, nullptr);
    std::memcpy(apos, S.pos.data(), sizeof(double) * S.pos.size());
    return n;
}

// Batch of CSR-packed systems relaxed on `nthreads` host threads, one system
// per task, dynamic schedule.  apos is updated in place; steps/E/Fmax out.
int orc_relax_batch(int nsys, const int* natoms, const int* nnode, double* apos, const double* REQs,
                    const int* neighs, const double* apars, const double* bLs, const double* bKs, const double* Ksp,
                    const double* Kpp, const int* sw, int gnx, int gny, int gnz, const double* gpos0,
                    const double* gstep, const double* gcoefs, double alpha, const double* fire, double Fconv,
                    int nmax, int conv_mode, int nthreads, int* steps_out, double* E_out, double* Fmax_out) {
    orc64::relax_batch(nsys, natoms, nnode, apos, REQs, neighs, apars, bLs, bKs, Ksp, Kpp, sw, gnx, gny, gnz, gpos0,
                       gstep, gcoefs, alpha, fire, Fconv, nmax, conv_mode, nthreads, steps_out, E_out, Fmax_out);
    return 0;
}

// The same relaxation with every array and every operation in single
// precision (the FP32 variant of the CPU baseline); inputs are converted once.
int orc_relax_batch_f32(int nsys, const int* natoms, const int* nnode, double* apos, const double* REQs,
                        const int* neighs, const double* apars, const double* bLs, const double* bKs,
                        const double* Ksp, const double* Kpp, const int* sw, int gnx, int gny, int gnz,
                        const double* gpos0, const double* gstep, const double* gcoefs, double alpha,
                        const double* fire, double Fconv, int nmax, int conv_mode, int nthreads, int* steps_out,
                        double* E_out, double* Fmax_out) {
    size_t nA = 0, nM = 0;
    for (int s = 0; s < nsys; s++) {
        nA += natoms[s];
        nM += nnode[s];
    }
    auto f32 = [](const double* x, size_t n) { return x ? std::vector<float>(x, x + n) : std::vector<float>(n, 0.f); };
    std::vector<float> pos = f32(apos, (nA + nM) * 3), req = f32(REQs, nA * 4), ap = f32(apars, nM * 4),
                       bl = f32(bLs, nM * 4), bk = f32(bKs, nM * 4), ks = f32(Ksp, nM * 4), kp = f32(Kpp, nM * 4),
                       g0 = f32(gpos0, 3), gs = f32(gstep, 3), fr = f32(fire, 8);
    std::vector<float> gc = gcoefs ? f32(gcoefs, (size_t)gnx * gny * gnz * 3) : std::vector<float>();  // P, L, Q per node
    std::vector<float> E(nsys), fm(nsys);
    orc32::relax_batch(nsys, natoms, nnode, pos.data(), req.data(), neighs, ap.data(), bl.data(), bk.data(), ks.data(),
                       kp.data(), sw, gnx, gny, gnz, gpos0 ? g0.data() : nullptr, gstep ? gs.data() : nullptr,
                       gcoefs ? gc.data() : nullptr, (float)alpha,
                       fr.data(), (float)Fconv, nmax, conv_mode, nthreads, steps_out, E.data(), fm.data());
    for (size_t i = 0; i < pos.size(); i++) apos[i] = pos[i];
    for (int s = 0; s < nsys; s++) {
        if (E_out) E_out[s] = E[s];
        if (Fmax_out) Fmax_out[s] = fm[s];
    }
    return 0;
}

// Term samplers with the reference harness semantics (libMMFF.h:439-473).
void orc_sample_terms(int kind, double k, double ang0, double r1, double r2, int n, const double* angles, double* Es,
                      double* Fs) {
    V h1{1, 0, 0}, f1, f2;
    for (int i = 0; i < n; i++) {
        V h2{std::cos(angles[i]), std::sin(angles[i]), 0.0};
        if (kind == 0) Es[i] = pi_align(h1, h2, 1. / r1, 1. / r2, k, f1, f2);
        else if (kind == 1) Es[i] = ang_cos(h1, h2, 1. / r1, 1. / r2, k, std::cos(ang0), f1, f2);
        else Es[i] = ang_cos_half(h1, h2, 1. / r1, 1. / r2, std::cos(ang0 / 2), std::sin(ang0 / 2), k, f1, f2);
        Fs[i] = f1.y;
    }
}
void orc_sample_pair(int n, const double* rs, const double* REQij, double* Es, double* Fs) {
    for (int i = 0; i < n; i++) {
        V f;
        Es[i] = pair_ljqh(V{rs[i], 0, 0}, REQij[0], REQij[1], REQij[2], REQij[3], f);
        Fs[i] = -f.x;
    }
}

}  // extern "C"
