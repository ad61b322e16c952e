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

namespace {

const double COULOMB_K = 14.3996448915;  // cpp/common/constants.h:9

struct V {
    double x, y, z;
};
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator*(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline void operator+=(V& a, V b) { a = a + b; }
inline void operator-=(V& a, V b) { a = a - b; }
inline double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V ld(const double* p) { return {p[0], p[1], p[2]}; }
inline void st(double* p, V v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

struct Sys {
    int N, M;
    std::vector<double> pos;  // (N+M)*3
    const double *REQ, *apars, *bLs, *bKs, *Ksp, *Kpp;
    const int* ng;
    int doBonds, doPiPiI, doPiSigma, doAngles, bNonBonded;
    // grid
    const double* gc = nullptr;
    int gn[3] = {0, 0, 0};
    double gp0[3], gstep[3], alpha = 1.5;
    // positional restraints and walls of the GPU update kernel (optional)
    const double* constr = nullptr;   // natoms x 4 (x, y, z, on if > 0)
    const double* constrK = nullptr;  // natoms x 4 (per-axis stiffness; negative = 0)
    double bbox[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // lower corner, upper corner, wall stiffness
    // periodic molecule (optional): image shifts and the image of each bonded neighbour
    int npbc = 0;
    const double* shifts = nullptr;  // npbc x 3
    const int* ngC = nullptr;        // natoms x 4
    // work
    std::vector<double> f, fng, fngpi;
};

// Explicit substrate atoms (instead of the grid) for the following calls
// (natoms = 0 clears); the arrays are copied.
struct Surf {
    int natoms = 0, npbc = 0;
    std::vector<double> apos, REQs, shifts;
    double K = -1.5;  // -alphaMorse
};
Surf g_surf;

// Periodic molecule for the following orc_eval / orc_relax / orc_explore calls
// on systems of `natoms` atoms (natoms = 0 clears); the arrays are copied.
struct Pbc {
    int natoms = 0, npbc = 0;
    std::vector<double> shifts;
    std::vector<int> ngC;
};
Pbc g_pbc;

// Restraints for the following orc_eval / orc_relax calls (natoms = 0 clears);
// the arrays are copied.
struct Restraints {
    int natoms = 0;
    std::vector<double> constr, constrK;
    double bbox[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};
Restraints g_restraints;

// getLJQH, Forces.h:341-360.  dp = pj - pi, returns E; f = force on i.
double pair_ljqh(V dp, double R, double E0, double Q, double H, V& f) {
    const double r2 = dot(dp, dp);
    const double ir2 = 1.0 / r2;
    double E = COULOMB_K * Q * std::sqrt(ir2);
    double F = E * ir2;
    const double u2 = R * R * ir2;
    const double u6 = u2 * u2 * u2;
    const double vdW = u6 * E0;
    const double Hb = u6 * u6 * ((H < 0) ? H * E0 : 0.0);
    E += (u6 - 2.) * vdW + Hb;
    F += ((u6 - 1.) * vdW + Hb) * ir2 * 12;
    f = dp * (-F);
    return E;
}

// evalAngleCosHalf, Forces.h:190-216
double ang_cos_half(V h1, V h2, double ir1, double ir2, double c0x, double c0y, double k, V& f1, V& f2) {
    V h = h1 + h2;
    double c2 = dot(h, h) * 0.25;
    double s2 = 1 - c2 + 1e-14;
    double c = std::sqrt(c2), s = std::sqrt(s2);
    double csx = c0x * c + c0y * s;  // cs0 udiv (c,s), Vec2.h:130
    double csy = c0y * c - c0x * s;
    double E = k * (1 - csx);
    double fr = -k * csy;
    c2 *= -2;
    fr /= 4 * c * s;
    double fr1 = fr * ir1, fr2 = fr * ir2;
    f1 = h * fr1 + h1 * (fr1 * c2);
    f2 = h * fr2 + h2 * (fr2 * c2);
    return E;
}
// evalAngleCos, Forces.h:218-228
double ang_cos(V h1, V h2, double ir1, double ir2, double K, double c0, V& f1, V& f2) {
    double c = dot(h1, h2);
    f1 = h2 - h1 * c;
    f2 = h1 - h2 * c;
    double c_ = c - c0;
    double E = K * c_ * c_;
    double fang = -K * c_ * 2;
    f1 = f1 * (fang * ir1);
    f2 = f2 * (fang * ir2);
    return E;
}
// evalPiAling, Forces.h:230-241
double pi_align(V h1, V h2, double ir1, double ir2, double K, V& f1, V& f2) {
    (void)ir1;
    (void)ir2;
    double c = dot(h1, h2);
    f1 = h2 - h1 * c;
    f2 = h1 - h2 * c;
    bool sign = c < 0;
    if (sign) c = -c;
    double E = K * (1 - c);
    double fang = sign ? -K : K;
    f1 = f1 * fang;
    f2 = f2 * fang;
    return E;
}

// eval_atom, MMFFsp3_loc.h:250-575 (fapos[ia] and fpipos[ia] are assigned)
double eval_node(Sys& S, int ia) {
    const int N = S.N, M = S.M;
    const V pa = ld(&S.pos[ia * 3]);
    const V hpi = ld(&S.pos[(N + ia) * 3]);
    const int* ings = S.ng + ia * 4;
    const double* bL = S.bLs + ia * 4;
    const double* bK = S.bKs + ia * 4;
    const double* ksp = S.Ksp + ia * 4;
    const double* kpp = S.Kpp + ia * 4;
    const double* apar = S.apars + ia * 4;
    V fa{0, 0, 0}, fpi{0, 0, 0};
    V fbs[4], fps[4], hs[4];
    double il[4];
    for (int i = 0; i < 4; i++) fbs[i] = fps[i] = V{0, 0, 0};
    double E = 0;
    V f1, f2;
    for (int i = 0; i < 4; i++) {
        int ing = ings[i];
        if (ing < 0) break;
        V h = ld(&S.pos[ing * 3]) - pa;
        if (S.npbc > 0) h += ld(&S.shifts[S.ngC[ia * 4 + i] * 3]);  // bond image, MMFFsp3_loc.h:325-329
        const double l = std::sqrt(dot(h, h));
        h = h * (1.0 / l);
        hs[i] = h;
        il[i] = 1.0 / l;
        if (ia < ing) {
            if (S.doBonds) {  // evalBond, Forces.h:137-143
                double dl = l - bL[i];
                double fr = dl * bK[i];
                f1 = h * fr;
                E += fr * dl * 0.5;
                fbs[i] -= f1;
                fa += f1;
            }
            double k = kpp[i];
            if (S.doPiPiI && ing < M && k > 1e-6) {
                E += pi_align(hpi, ld(&S.pos[(N + ing) * 3]), 1., 1., k, f1, f2);
                fpi += f1;
                fps[i] += f2;
            }
        }
        double k = ksp[i];
        if (S.doPiSigma && k > 1e-6) {
            E += ang_cos(hpi, h, 1., il[i], k, apar[3], f1, f2);
            fpi += f1;
            fa -= f2;
            fbs[i] += f2;
        }
    }
    if (S.doAngles) {
        const double ssK = apar[2];
        for (int i = 0; i < 3; i++) {
            int ing = ings[i];
            if (ing < 0) break;
            for (int j = i + 1; j < 4; j++) {
                int jng = ings[j];
                if (jng < 0) break;
                E += ang_cos_half(hs[i], hs[j], il[i], il[j], apar[0], apar[1], ssK, f1, f2);
                fa -= f1 + f2;
                if (S.bNonBonded) {  // bSubtractAngleNonBond
                    const double* Ri = S.REQ + ing * 4;
                    const double* Rj = S.REQ + jng * 4;
                    V dp = hs[j] * (1.0 / il[j]) - hs[i] * (1.0 / il[i]);
                    V fij;
                    E -= pair_ljqh(dp, Ri[0] + Rj[0], Ri[1] * Rj[1], Ri[2] * Rj[2], Ri[3] * Rj[3], fij);
                    f1 -= fij;
                    f2 += fij;
                }
                fbs[i] += f1;
                fbs[j] += f2;
            }
        }
    }
    for (int i = 0; i < 4; i++) {
        st(&S.fng[(ia * 4 + i) * 3], fbs[i]);
        st(&S.fngpi[(ia * 4 + i) * 3], fps[i]);
    }
    st(&S.f[ia * 3], fa);
    st(&S.f[(N + ia) * 3], fpi);
    return E;
}

// evalLJQs_ng4_PBC_atom_omp, NBFF.h:575-620 (adds to fapos[ia], returns
// 0.5*E): every other atom at every image shift; a bonded neighbour is skipped
// only at the image of the bond (neighCell).
double nb_atom_pbc(Sys& S, int ia) {
    const V pi = ld(&S.pos[ia * 3]);
    const double* Ri = S.REQ + ia * 4;
    const int* ng = S.ng + ia * 4;
    const int* ngC = S.ngC + ia * 4;
    double E = 0;
    V fi{0, 0, 0};
    for (int j = 0; j < S.N; j++) {
        if (j == ia) continue;
        const double* Rj = S.REQ + j * 4;
        const V dp = ld(&S.pos[j * 3]) - pi;
        for (int ipbc = 0; ipbc < S.npbc; ipbc++) {
            bool skip = false;
            for (int k = 0; k < 4; k++) skip |= (j == ng[k]) && (ipbc == ngC[k]);
            if (skip) continue;
            V fij;
            E += pair_ljqh(dp + ld(&S.shifts[ipbc * 3]), Ri[0] + Rj[0], Ri[1] * Rj[1], Ri[2] * Rj[2],
                           Ri[3] * Rj[3], fij);
            fi += fij;
        }
    }
    st(&S.f[ia * 3], ld(&S.f[ia * 3]) + fi);
    return 0.5 * E;
}

// evalLJQs_ng4_atom_omp, NBFF.h:666-691 (adds to fapos[ia], returns 0.5*E)
double nb_atom(Sys& S, int ia) {
    if (S.npbc > 0) return nb_atom_pbc(S, ia);
    const V pi = ld(&S.pos[ia * 3]);
    const double* Ri = S.REQ + ia * 4;
    const int* ng = S.ng + ia * 4;
    double E = 0;
    V fi{0, 0, 0};
    for (int j = 0; j < S.N; j++) {
        if (j == ia || j == ng[0] || j == ng[1] || j == ng[2] || j == ng[3]) continue;
        const double* Rj = S.REQ + j * 4;
        V fij;
        E += pair_ljqh(ld(&S.pos[j * 3]) - pi, Ri[0] + Rj[0], Ri[1] * Rj[1], Ri[2] * Rj[2], Ri[3] * Rj[3], fij);
        fi += fij;
    }
    V f = ld(&S.f[ia * 3]) + fi;
    st(&S.f[ia * 3], f);
    return 0.5 * E;
}

// cubic B-spline basis, Bspline.h:253-276
void basis(double u, double* b) {
    const double inv6 = 1. / 6.;
    const double u2 = u * u, t = 1 - u;
    b[0] = inv6 * t * t * t;
    b[1] = inv6 * (3 * u2 * (u - 2) + 4);
    b[2] = inv6 * (3 * u * (1 + u - u2) + 1);
    b[3] = inv6 * u2 * u;
}
void dbasis(double u, double* d) {
    const double u2 = u * u, t = 1 - u;
    d[0] = -0.5 * t * t;
    d[1] = 0.5 * (3 * u2 - 4 * u);
    d[2] = 0.5 * (-3 * u2 + 2 * u + 1);
    d[3] = 0.5 * u2;
}
int modulo(int i, int m) {  // Bspline.h:83-89
    int r = i % m;
    if (r < 0) r += m;
    return r;
}
// periodic tap offsets: choose_inds_pbc over make_inds_pbc tables, Bspline.h:39-60
void tap_offsets(int i, int n, int* o) {
    if (i >= n - 3) {
        const int ii = i + 4 - n;
        for (int k = 0; k < 4; k++) o[k] = (k < 4 - ii) ? k : k - n;
    } else {
        for (int k = 0; k < 4; k++) o[k] = k;
    }
}

// getForce_Bspline + fe3d_pbc_comb3 (GridFF.h:355-402, Bspline.h:742-809)
double grid_sample(const Sys& S, V p, const double* PLQ, V& fout) {
    const int nx = S.gn[0], ny = S.gn[1], nz = S.gn[2];
    const double ux = (p.x - S.gp0[0]) / S.gstep[0];
    const double uy = (p.y - S.gp0[1]) / S.gstep[1];
    const double uz = (p.z - S.gp0[2]) / S.gstep[2];
    int ix = (int)ux, iy = (int)uy, iz = (int)uz;
    if (ux < 0) ix--;
    if (uy < 0) iy--;
    const double tx = ux - ix, ty = uy - iy, tz = uz - iz;
    iz = std::min(std::max(iz, 3), nz - 5);
    ix = modulo(ix - 1, nx);
    iy = modulo(iy - 1, ny);
    int qx[4], qy[4];
    tap_offsets(ix, nx, qx);
    tap_offsets(iy, ny, qy);
    double bz[4], dz[4], by[4], dy[4], bx[4], dx[4];
    basis(tz, bz);
    dbasis(tz, dz);
    basis(ty, by);
    dbasis(ty, dy);
    basis(tx, bx);
    dbasis(tx, dx);
    const int i0 = (iz - 1) + nz * (iy + ny * ix);
    double Ex[4][3];  // per x tap: (Fy, Fz, E)  (fe2d_comb3)
    for (int a = 0; a < 4; a++) {
        double fe[4][2];
        for (int b = 0; b < 4; b++) {
            const double* E = S.gc + 3 * (size_t)(i0 + qx[a] * ny * nz + qy[b] * nz);
            double cs[4];
            for (int c = 0; c < 4; c++)
                cs[c] = PLQ[0] * E[c * 3] + PLQ[1] * E[c * 3 + 1] + PLQ[2] * E[c * 3 + 2];
            fe[b][0] = bz[0] * cs[0] + bz[1] * cs[1] + bz[2] * cs[2] + bz[3] * cs[3];
            fe[b][1] = dz[0] * cs[0] + dz[1] * cs[1] + dz[2] * cs[2] + dz[3] * cs[3];
        }
        Ex[a][0] = fe[0][0] * dy[0] + fe[1][0] * dy[1] + fe[2][0] * dy[2] + fe[3][0] * dy[3];
        Ex[a][1] = fe[0][1] * by[0] + fe[1][1] * by[1] + fe[2][1] * by[2] + fe[3][1] * by[3];
        Ex[a][2] = fe[0][0] * by[0] + fe[1][0] * by[1] + fe[2][0] * by[2] + fe[3][0] * by[3];
    }
    const double Fx = dx[0] * Ex[0][2] + dx[1] * Ex[1][2] + dx[2] * Ex[2][2] + dx[3] * Ex[3][2];
    const double Fy = bx[0] * Ex[0][0] + bx[1] * Ex[1][0] + bx[2] * Ex[2][0] + bx[3] * Ex[3][0];
    const double Fz = bx[0] * Ex[0][1] + bx[1] * Ex[1][1] + bx[2] * Ex[2][1] + bx[3] * Ex[3][1];
    const double E = bx[0] * Ex[0][2] + bx[1] * Ex[1][2] + bx[2] * Ex[2][2] + bx[3] * Ex[3][2];
    fout += V{-Fx / S.gstep[0], -Fy / S.gstep[1], -Fz / S.gstep[2]};
    return E;
}

// REQ2PLQ_d, Forces.h:611-619
void req2plq(const double* REQ, double alpha, double* PLQ) {
    double e = std::exp(alpha * REQ[0]);
    double cL = e * REQ[1];
    PLQ[0] = e * cL;
    PLQ[1] = cL;
    PLQ[2] = REQ[2];
    PLQ[3] = e * e * REQ[3];
}

// Substrate atoms, GridFF::evalMorsePBC (GridFF.h:1229-1261) with
// addAtomicForceMorseQ (Forces.h:540-561) and combineREQ (Forces.h:263-268);
// the pair's Morse amplitude E0*e is rounded to float as the reference does
// (`float Ae`).  Adds the force to fapos[ia], returns the energy.
double surf_atom(Sys& S, int ia) {
    const V pi = ld(&S.pos[ia * 3]);
    const double* Ri = S.REQ + ia * 4;
    const double K = g_surf.K;
    double E = 0;
    V f{0, 0, 0};
    for (int j = 0; j < g_surf.natoms; j++) {
        const double* Rj = &g_surf.REQs[j * 4];
        const double r0 = Rj[0] + Ri[0], E0 = Rj[1] * Ri[1], qq = Rj[2] * Ri[2];
        const V dp0 = pi - ld(&g_surf.apos[j * 3]);
        for (int ipbc = 0; ipbc < g_surf.npbc; ipbc++) {
            const V dp = dp0 + ld(&g_surf.shifts[ipbc * 3]);
            const double r2 = dot(dp, dp);
            const double ir2 = 1 / r2;
            const double r = std::sqrt(r2 + 1e-32);
            const double ir = std::sqrt(ir2);
            const double e = std::exp(K * (r - r0));
            const float Ae = (float)(E0 * e);
            const double fM = Ae * 2 * K * (e - 1);
            const double EM = Ae * (e - 2);
            const double Eel = COULOMB_K * qq * ir;
            f -= dp * (fM / r - Eel * ir2);
            E += EM + Eel;
        }
    }
    st(&S.f[ia * 3], ld(&S.f[ia * 3]) + f);
    return E;
}

// eval_no_omp (MolWorld_sp3.h:2178-2245) + assemble_atom (MMFFsp3_loc.h:998-1021)
double eval_sys(Sys& S) {
    const int N = S.N, M = S.M;
    S.f.assign((size_t)(N + M) * 3, 0.0);
    S.fng.assign((size_t)M * 12, 0.0);
    S.fngpi.assign((size_t)M * 12, 0.0);
    double E = 0;
    for (int ia = 0; ia < N; ia++) {
        st(&S.f[ia * 3], V{0, 0, 0});
        if (ia < M) st(&S.f[(N + ia) * 3], V{0, 0, 0});
        if (ia < M) E += eval_node(S, ia);
        if (S.bNonBonded) E += nb_atom(S, ia);
        if (S.gc) {
            double plq[4];
            req2plq(S.REQ + ia * 4, S.alpha, plq);
            V f = ld(&S.f[ia * 3]);
            E += grid_sample(S, ld(&S.pos[ia * 3]), plq, f);
            st(&S.f[ia * 3], f);
        } else if (g_surf.natoms > 0) {  // bSurfAtoms without bGridFF (MolWorld_sp3.h:2225-2229)
            E += surf_atom(S, ia);
        }
    }
    // back-neighbours as in makeBackNeighs (first-empty fill over nodes)
    std::vector<int> bk((size_t)N * 4, -1);
    for (int ia = 0; ia < M; ia++)
        for (int j = 0; j < 4; j++) {
            int ja = S.ng[ia * 4 + j];
            if (ja < 0) continue;
            int* row = &bk[ja * 4];
            int k = 0;
            while (k < 4 && row[k] >= 0) k++;
            if (k < 4) row[k] = ia * 4 + j;
        }
    for (int ia = 0; ia < N; ia++) {
        V fa{0, 0, 0}, fp{0, 0, 0};
        const bool bpi = ia < M;
        for (int i = 0; i < 4; i++) {
            int j = bk[ia * 4 + i];
            if (j < 0) break;
            fa += ld(&S.fng[j * 3]);
            if (bpi) fp += ld(&S.fngpi[j * 3]);
        }
        st(&S.f[ia * 3], ld(&S.f[ia * 3]) + fa);
        // restraint and wall of the GPU update kernel, updateAtomsMMFFf4
        // (relax_multi.cl:997-1009), as part of the atom force
        if (S.constr && S.constr[ia * 4 + 3] > 0) {
            const double* c = S.constr + ia * 4;
            const double* k = S.constrK + ia * 4;
            const V p = ld(&S.pos[ia * 3]);
            const V fc{(c[0] - p.x) * std::max(k[0], 0.0), (c[1] - p.y) * std::max(k[1], 0.0),
                       (c[2] - p.z) * std::max(k[2], 0.0)};
            st(&S.f[ia * 3], ld(&S.f[ia * 3]) + fc);
        }
        for (int d = 0; d < 3; d++) {
            const double kw = S.bbox[6 + d];
            if (!(kw > 0)) continue;
            const double x = S.pos[ia * 3 + d], lo = S.bbox[d], hi = S.bbox[3 + d];
            if (x < lo) S.f[ia * 3 + d] += (lo - x) * kw;
            else if (x > hi) S.f[ia * 3 + d] += (hi - x) * kw;
        }
        if (bpi) {
            V f = ld(&S.f[(N + ia) * 3]) + fp;
            V p = ld(&S.pos[(N + ia) * 3]);
            f -= p * dot(p, f);  // makeOrthoU, Vec3.h:127
            st(&S.f[(N + ia) * 3], f);
        }
    }
    return E;
}

void setup(Sys& S, int natoms, int nnode, const double* apos, const double* REQs, const int* neighs,
           const double* apars, const double* bLs, const double* bKs, const double* Ksp, const double* Kpp,
           const int* sw, int gnx, int gny, int gnz, const double* gpos0, const double* gstep, const double* gcoefs,
           double alpha) {
    S.N = natoms;
    S.M = nnode;
    S.pos.assign(apos, apos + (size_t)(natoms + nnode) * 3);
    S.REQ = REQs;
    S.ng = neighs;
    S.apars = apars;
    S.bLs = bLs;
    S.bKs = bKs;
    S.Ksp = Ksp;
    S.Kpp = Kpp;
    S.doBonds = sw[0];
    S.doPiPiI = sw[1];
    S.doPiSigma = sw[2];
    S.doAngles = sw[3];
    S.bNonBonded = sw[4];
    if (g_pbc.natoms == natoms && g_pbc.npbc > 0) {
        S.npbc = g_pbc.npbc;
        S.shifts = g_pbc.shifts.data();
        S.ngC = g_pbc.ngC.data();
    }
    if (g_restraints.natoms == natoms) {
        S.constr = g_restraints.constr.data();
        S.constrK = g_restraints.constrK.data();
        for (int k = 0; k < 9; k++) S.bbox[k] = g_restraints.bbox[k];
    }
    S.gc = gcoefs;
    if (gcoefs) {
        S.gn[0] = gnx;
        S.gn[1] = gny;
        S.gn[2] = gnz;
        for (int k = 0; k < 3; k++) {
            S.gp0[k] = gpos0[k];
            S.gstep[k] = gstep[k];
        }
        S.alpha = alpha;
    }
}

struct Fire {
    double dt, dt_max, dt_min, damping, damp_max, finc, fdec, falpha;
    int minLastNeg, lastNeg;
    double cv, cf, renorm;
};

// FIRE_update_params (DynamicOpt.cpp:250-267), damp_func_FIRE (:210-222)
void fire_update(Fire& F, double vf, double vv, double ff) {
    const double f_len = std::sqrt(ff), v_len = std::sqrt(vv);
    const double cos_vf = vf / (f_len * v_len + 1e-32);
    F.renorm = v_len / (f_len + 1e-32);
    double cf;
    if (cos_vf < 0.0) {
        F.cv = 0.0;  // cv_kill
        cf = 0.0;
    } else {
        F.cv = 1 - F.damping;
        cf = F.damping;
    }
    F.cf = F.renorm * cf;
    if (cos_vf < 0.0) {
        F.dt = std::max(F.dt * F.fdec, F.dt_min);
        F.lastNeg = 0;
        F.damping = F.damp_max;
    } else {
        if (F.lastNeg > F.minLastNeg) {
            F.dt = std::min(F.dt * F.finc, F.dt_max);
            F.damping = F.damping * F.falpha;
        }
        F.lastNeg++;
    }
}

// fire = {dt_max, dt_min, damp_max, finc, fdec, falpha, minLastNeg, Flimit}
int relax_sys(Sys& S, const double* fire, double Fconv, int nmax, int conv_mode, double* E_out, double* traj,
              int ntraj, double* fmax_out) {
    const double Flimit = fire[7];
    const int N = S.N, V_ = S.N + S.M;
    Fire F{};
    F.dt = F.dt_max = fire[0];
    F.dt_min = fire[1];
    F.damping = F.damp_max = fire[2];
    F.finc = fire[3];
    F.fdec = fire[4];
    F.falpha = fire[5];
    F.minLastNeg = (int)fire[6];
    F.lastNeg = 0;
    std::vector<double> vel((size_t)V_ * 3, 0.0);
    const double F2conv = Fconv * Fconv;
    int itr = 0;
    double E = 0, fm2 = 0;
    for (itr = 0; itr < nmax; itr++) {
        E = eval_sys(S);
        double vf = 0, vv = 0, ff = 0;
        fm2 = 0;
        for (int i = 0; i < V_; i++) fm2 = std::max(fm2, dot(ld(&S.f[i * 3]), ld(&S.f[i * 3])));
        if (Flimit > 0) {  // per-vec clamp of the GPU update kernel, relax_multi.cl:980-983
            for (int i = 0; i < V_; i++) {
                V f = ld(&S.f[i * 3]);
                double f2 = dot(f, f);
                if (f2 > Flimit * Flimit) st(&S.f[i * 3], f * (Flimit / std::sqrt(f2)));
            }
        }
        for (int i = 0; i < V_ * 3; i++) {
            vf += vel[i] * S.f[i];
            vv += vel[i] * vel[i];
            ff += S.f[i] * S.f[i];
        }
        if (!std::isfinite(ff)) {  // blown up: report a non-finite Fmax
            fm2 = NAN;
            break;
        }
        if (conv_mode == 0 && fm2 < F2conv) break;
        fire_update(F, vf, vv, ff);
        // move_atom_FIRE receives dt as float (MMFFsp3_loc.h:1264)
        const double dt = (double)(float)F.dt;
        for (int i = 0; i < V_; i++) {
            V f = ld(&S.f[i * 3]);
            V v = ld(&vel[i * 3]);
            V p = ld(&S.pos[i * 3]);
            const bool bPi = i >= N;
            if (bPi) f -= p * dot(p, f);
            v = v * F.cv;
            v += f * (dt + F.renorm * F.cf);
            if (bPi) v -= p * dot(p, v);
            p += v * dt;
            if (bPi) p = p * (1.0 / std::sqrt(dot(p, p)));
            st(&vel[i * 3], v);
            st(&S.pos[i * 3], p);
        }
        if (traj && itr < ntraj) std::memcpy(traj + (size_t)itr * V_ * 3, S.pos.data(), sizeof(double) * V_ * 3);
        if (conv_mode == 1 && ff < F2conv) break;
    }
    if (E_out) *E_out = E;
    if (fmax_out) *fmax_out = std::sqrt(fm2);
    return itr;
}

// Random number of the exploring drive: the counter-based hash of (seed,
// system, vec, axis, step) the device draws (murmur3 32-bit finaliser after
// each field), as a uniform on [-sqrt(3), sqrt(3)).
uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}
double drive_xi(uint32_t seed, uint32_t sys, uint32_t vec, uint32_t axis, uint32_t step) {
    uint32_t h = fmix32(seed + 0x9E3779B9u);
    h = fmix32(h ^ sys);
    h = fmix32(h ^ (vec * 3u + axis));
    h = fmix32(h ^ step) >> 8;
    return ((double)h * (2.0 / 16777216.0) - 1.0) * std::sqrt(3.0);
}

// Exploring (global optimisation) of one system: the phase machine of
// MolWorld_sp3_multi::updateMultiExploring (MolWorld_sp3_multi.h:770-833) per
// step, with the Langevin drive of updateAtomsMMFFf4 (relax_multi.cl:
// 1160-1178: f += -gamma*v + xi*sqrt(2*kB*T*gamma/dt) on atoms) and the
// exploring MD parameters of evalVF_new (:755-760: dt = dt_max, velocity kept,
// no FIRE mixing).  Relaxing phases run FIRE as relax_sys; a converged one
// records a minimum (energy, positions) and starts an exploring phase of
// nExplore steps (so does a relaxation longer than nRelax > 0 steps, without a
// record); after it FIRE restarts from dt_max, damp_max.  Runs steps 0..nmax.
int explore_sys(Sys& S, const double* fire, double Fconv, int nmax, double T, double gamma, int nExplore, int nRelax,
                uint32_t seed, uint32_t sysidx, int kmax, double* minE, double* minpos, double* traj, int ntraj) {
    const double kB = 8.617333262145e-5;  // eV/K, cpp/common/constants.h:12
    const double Flimit = fire[7];
    const int N = S.N, V_ = S.N + S.M;
    Fire F{};
    F.dt_max = fire[0];
    F.dt_min = fire[1];
    F.damp_max = fire[2];
    F.finc = fire[3];
    F.fdec = fire[4];
    F.falpha = fire[5];
    F.minLastNeg = (int)fire[6];
    auto fire_reset = [&] {
        F.dt = F.dt_max;
        F.damping = F.damp_max;
        F.lastNeg = 0;
    };
    fire_reset();
    std::vector<double> vel((size_t)V_ * 3, 0.0);
    const double F2conv = Fconv * Fconv;
    int phase = 0, pstep = 0, nfound = 0;
    for (int step = 0; step <= nmax; step++) {
        const double E = eval_sys(S);
        double fm2 = 0, vf = 0, vv = 0, ff = 0;
        for (int i = 0; i < V_; i++) fm2 = std::max(fm2, dot(ld(&S.f[i * 3]), ld(&S.f[i * 3])));
        if (Flimit > 0)
            for (int i = 0; i < V_; i++) {
                V f = ld(&S.f[i * 3]);
                double f2 = dot(f, f);
                if (f2 > Flimit * Flimit) st(&S.f[i * 3], f * (Flimit / std::sqrt(f2)));
            }
        for (int i = 0; i < V_ * 3; i++) {
            vf += vel[i] * S.f[i];
            vv += vel[i] * vel[i];
            ff += S.f[i] * S.f[i];
        }
        if (!std::isfinite(ff)) return -1;
        bool reset = false;
        if (phase == 0) {
            const bool conv = fm2 < F2conv;
            if (conv || (nRelax > 0 && pstep >= nRelax)) {
                if (conv) {
                    if (nfound < kmax) {
                        if (minE) minE[nfound] = E;
                        if (minpos) std::memcpy(minpos + (size_t)nfound * V_ * 3, S.pos.data(), sizeof(double) * V_ * 3);
                    }
                    nfound++;
                }
                phase = 1;
                pstep = 0;
            }
        } else if (pstep >= nExplore) {
            phase = 0;
            pstep = 0;
            reset = true;
        }
        if (step >= nmax) break;
        const bool drive = phase == 1;
        double dt, cv, cfe;
        if (drive) {
            dt = (double)(float)F.dt_max;
            cv = 1.0;
            cfe = 0.0;
        } else {
            if (reset) fire_reset();
            fire_update(F, vf, vv, ff);
            dt = (double)(float)F.dt;
            cv = F.cv;
            cfe = F.renorm * F.cf;
        }
        const double amp = drive ? std::sqrt(2.0 * kB * std::max(T, 0.0) * std::max(gamma, 0.0) / dt) : 0.0;
        for (int i = 0; i < V_; i++) {
            V f = ld(&S.f[i * 3]);
            V v = ld(&vel[i * 3]);
            V p = ld(&S.pos[i * 3]);
            const bool bPi = i >= N;
            if (bPi) {
                f -= p * dot(p, f);
            } else if (drive) {
                f -= v * std::max(gamma, 0.0);
                f += V{amp * drive_xi(seed, sysidx, (uint32_t)i, 0u, (uint32_t)step),
                       amp * drive_xi(seed, sysidx, (uint32_t)i, 1u, (uint32_t)step),
                       amp * drive_xi(seed, sysidx, (uint32_t)i, 2u, (uint32_t)step)};
            }
            v = v * cv;
            v += f * (dt + cfe);
            if (bPi) v -= p * dot(p, v);
            p += v * dt;
            if (bPi) p = p * (1.0 / std::sqrt(dot(p, p)));
            st(&vel[i * 3], v);
            st(&S.pos[i * 3], p);
        }
        pstep++;
        if (traj && step < ntraj) std::memcpy(traj + (size_t)step * V_ * 3, S.pos.data(), sizeof(double) * V_ * 3);
    }
    return nfound;
}

}  // namespace

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
    std::vector<size_t> a0(nsys + 1, 0), n0(nsys + 1, 0), v0(nsys + 1, 0);
    for (int s = 0; s < nsys; s++) {
        a0[s + 1] = a0[s] + natoms[s];
        n0[s + 1] = n0[s] + nnode[s];
        v0[s + 1] = v0[s] + natoms[s] + nnode[s];
    }
    std::atomic<int> next{0};
    auto worker = [&]() {
        for (;;) {
            int s = next.fetch_add(1);
            if (s >= nsys) break;
            Sys S;
            setup(S, natoms[s], nnode[s], apos + v0[s] * 3, REQs + a0[s] * 4, neighs + a0[s] * 4, apars + n0[s] * 4,
                  bLs + n0[s] * 4, bKs + n0[s] * 4, Ksp + n0[s] * 4, Kpp + n0[s] * 4, sw, gnx, gny, gnz, gpos0, gstep,
                  gcoefs, alpha);
            double E = 0, fm = 0;
            int n = relax_sys(S, fire, Fconv, nmax, conv_mode, &E, nullptr, 0, &fm);
            std::memcpy(apos + v0[s] * 3, S.pos.data(), sizeof(double) * S.pos.size());
            if (steps_out) steps_out[s] = n;
            if (E_out) E_out[s] = E;
            if (Fmax_out) Fmax_out[s] = fm;
        }
    };
    nthreads = std::max(1, nthreads);
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(worker);
    for (auto& t : th) t.join();
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
