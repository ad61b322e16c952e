// Core of the CPU oracle, written for a scalar type `real` that the including
// file defines before including it (double for the oracle proper, float for
// the FP32 variant of the CPU baseline).  TEST INFRASTRUCTURE ONLY: see
// oracle.cpp for what each function restates.
// no include guard: included once per scalar type, inside its own namespace
// small per-pair / per-atom helpers forced inline: a call passing the 3-float
// vector struct by value costs more than the helper itself in the float build
#ifndef ORC_HOT
#define ORC_HOT __attribute__((always_inline)) inline
#endif

const real COULOMB_K = real(14.3996448915);  // cpp/common/constants.h:9

struct V {
    real x, y, z;
};
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator*(V a, real s) { return {a.x * s, a.y * s, a.z * s}; }
inline void operator+=(V& a, V b) { a = a + b; }
inline void operator-=(V& a, V b) { a = a - b; }
inline real dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V ld(const real* p) { return {p[0], p[1], p[2]}; }
inline void st(real* p, V v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

struct Sys {
    int N, M;
    std::vector<real> pos;  // (N+M)*3
    const real *REQ, *apars, *bLs, *bKs, *Ksp, *Kpp;
    const int* ng;
    int doBonds, doPiPiI, doPiSigma, doAngles, bNonBonded;
    // grid
    const real* gc = nullptr;
    int gn[3] = {0, 0, 0};
    real gp0[3], gstep[3], alpha = real(1.5);
    // positional restraints and walls of the GPU update kernel (optional)
    const real* constr = nullptr;   // natoms x 4 (x, y, z, on if > 0)
    const real* constrK = nullptr;  // natoms x 4 (per-axis stiffness; negative = 0)
    real bbox[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // lower corner, upper corner, wall stiffness
    // periodic molecule (optional): image shifts and the image of each bonded neighbour
    int npbc = 0;
    const real* shifts = nullptr;  // npbc x 3
    const int* ngC = nullptr;        // natoms x 4
    // work
    std::vector<real> f, fng, fngpi;
};

// Explicit substrate atoms (instead of the grid) for the following calls
// (natoms = 0 clears); the arrays are copied.
struct Surf {
    int natoms = 0, npbc = 0;
    std::vector<real> apos, REQs, shifts;
    real K = -real(1.5);  // -alphaMorse
};
Surf g_surf;

// Periodic molecule for the following orc_eval / orc_relax / orc_explore calls
// on systems of `natoms` atoms (natoms = 0 clears); the arrays are copied.
struct Pbc {
    int natoms = 0, npbc = 0;
    std::vector<real> shifts;
    std::vector<int> ngC;
};
Pbc g_pbc;

// Restraints for the following orc_eval / orc_relax calls (natoms = 0 clears);
// the arrays are copied.
struct Restraints {
    int natoms = 0;
    std::vector<real> constr, constrK;
    real bbox[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};
Restraints g_restraints;

// getLJQH, Forces.h:341-360.  dp = pj - pi, returns E; f = force on i.
ORC_HOT real pair_ljqh(V dp, real R, real E0, real Q, real H, V& f) {
    const real r2 = dot(dp, dp);
    const real ir2 = real(1.0) / r2;
    real E = COULOMB_K * Q * std::sqrt(ir2);
    real F = E * ir2;
    const real u2 = R * R * ir2;
    const real u6 = u2 * u2 * u2;
    const real vdW = u6 * E0;
    const real Hb = u6 * u6 * ((H < 0) ? H * E0 : real(0.0));
    E += (u6 - real(2.)) * vdW + Hb;
    F += ((u6 - real(1.)) * vdW + Hb) * ir2 * 12;
    f = dp * (-F);
    return E;
}

// evalAngleCosHalf, Forces.h:190-216
ORC_HOT real ang_cos_half(V h1, V h2, real ir1, real ir2, real c0x, real c0y, real k, V& f1, V& f2) {
    V h = h1 + h2;
    real c2 = dot(h, h) * real(0.25);
    real s2 = 1 - c2 + real(1e-14);
    real c = std::sqrt(c2), s = std::sqrt(s2);
    real csx = c0x * c + c0y * s;  // cs0 udiv (c,s), Vec2.h:130
    real csy = c0y * c - c0x * s;
    real E = k * (1 - csx);
    real fr = -k * csy;
    c2 *= -2;
    fr /= 4 * c * s;
    real fr1 = fr * ir1, fr2 = fr * ir2;
    f1 = h * fr1 + h1 * (fr1 * c2);
    f2 = h * fr2 + h2 * (fr2 * c2);
    return E;
}
// evalAngleCos, Forces.h:218-228
ORC_HOT real ang_cos(V h1, V h2, real ir1, real ir2, real K, real c0, V& f1, V& f2) {
    real c = dot(h1, h2);
    f1 = h2 - h1 * c;
    f2 = h1 - h2 * c;
    real c_ = c - c0;
    real E = K * c_ * c_;
    real fang = -K * c_ * 2;
    f1 = f1 * (fang * ir1);
    f2 = f2 * (fang * ir2);
    return E;
}
// evalPiAling, Forces.h:230-241
ORC_HOT real pi_align(V h1, V h2, real ir1, real ir2, real K, V& f1, V& f2) {
    (void)ir1;
    (void)ir2;
    real c = dot(h1, h2);
    f1 = h2 - h1 * c;
    f2 = h1 - h2 * c;
    bool sign = c < 0;
    if (sign) c = -c;
    real E = K * (1 - c);
    real fang = sign ? -K : K;
    f1 = f1 * fang;
    f2 = f2 * fang;
    return E;
}

// eval_atom, MMFFsp3_loc.h:250-575 (fapos[ia] and fpipos[ia] are assigned)
real eval_node(Sys& S, int ia) {
    const int N = S.N, M = S.M;
    const V pa = ld(&S.pos[ia * 3]);
    const V hpi = ld(&S.pos[(N + ia) * 3]);
    const int* ings = S.ng + ia * 4;
    const real* bL = S.bLs + ia * 4;
    const real* bK = S.bKs + ia * 4;
    const real* ksp = S.Ksp + ia * 4;
    const real* kpp = S.Kpp + ia * 4;
    const real* apar = S.apars + ia * 4;
    V fa{0, 0, 0}, fpi{0, 0, 0};
    V fbs[4], fps[4], hs[4];
    real il[4];
    for (int i = 0; i < 4; i++) fbs[i] = fps[i] = V{0, 0, 0};
    real E = 0;
    V f1, f2;
    for (int i = 0; i < 4; i++) {
        int ing = ings[i];
        if (ing < 0) break;
        V h = ld(&S.pos[ing * 3]) - pa;
        if (S.npbc > 0) h += ld(&S.shifts[S.ngC[ia * 4 + i] * 3]);  // bond image, MMFFsp3_loc.h:325-329
        const real l = std::sqrt(dot(h, h));
        h = h * (real(1.0) / l);
        hs[i] = h;
        il[i] = real(1.0) / l;
        if (ia < ing) {
            if (S.doBonds) {  // evalBond, Forces.h:137-143
                real dl = l - bL[i];
                real fr = dl * bK[i];
                f1 = h * fr;
                E += fr * dl * real(0.5);
                fbs[i] -= f1;
                fa += f1;
            }
            real k = kpp[i];
            if (S.doPiPiI && ing < M && k > real(1e-6)) {
                E += pi_align(hpi, ld(&S.pos[(N + ing) * 3]), real(1.), real(1.), k, f1, f2);
                fpi += f1;
                fps[i] += f2;
            }
        }
        real k = ksp[i];
        if (S.doPiSigma && k > real(1e-6)) {
            E += ang_cos(hpi, h, real(1.), il[i], k, apar[3], f1, f2);
            fpi += f1;
            fa -= f2;
            fbs[i] += f2;
        }
    }
    if (S.doAngles) {
        const real ssK = apar[2];
        for (int i = 0; i < 3; i++) {
            int ing = ings[i];
            if (ing < 0) break;
            for (int j = i + 1; j < 4; j++) {
                int jng = ings[j];
                if (jng < 0) break;
                E += ang_cos_half(hs[i], hs[j], il[i], il[j], apar[0], apar[1], ssK, f1, f2);
                fa -= f1 + f2;
                if (S.bNonBonded) {  // bSubtractAngleNonBond
                    const real* Ri = S.REQ + ing * 4;
                    const real* Rj = S.REQ + jng * 4;
                    V dp = hs[j] * (real(1.0) / il[j]) - hs[i] * (real(1.0) / il[i]);
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
real nb_atom_pbc(Sys& S, int ia) {
    const V pi = ld(&S.pos[ia * 3]);
    const real* Ri = S.REQ + ia * 4;
    const int* ng = S.ng + ia * 4;
    const int* ngC = S.ngC + ia * 4;
    real E = 0;
    V fi{0, 0, 0};
    for (int j = 0; j < S.N; j++) {
        if (j == ia) continue;
        const real* Rj = S.REQ + j * 4;
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
    return real(0.5) * E;
}

// evalLJQs_ng4_atom_omp, NBFF.h:666-691 (adds to fapos[ia], returns 0.5*E)
real nb_atom(Sys& S, int ia) {
    if (S.npbc > 0) return nb_atom_pbc(S, ia);
    const V pi = ld(&S.pos[ia * 3]);
    const real* Ri = S.REQ + ia * 4;
    const int* ng = S.ng + ia * 4;
    real E = 0;
    V fi{0, 0, 0};
    for (int j = 0; j < S.N; j++) {
        if (j == ia || j == ng[0] || j == ng[1] || j == ng[2] || j == ng[3]) continue;
        const real* Rj = S.REQ + j * 4;
        V fij;
        E += pair_ljqh(ld(&S.pos[j * 3]) - pi, Ri[0] + Rj[0], Ri[1] * Rj[1], Ri[2] * Rj[2], Ri[3] * Rj[3], fij);
        fi += fij;
    }
    V f = ld(&S.f[ia * 3]) + fi;
    st(&S.f[ia * 3], f);
    return real(0.5) * E;
}

// cubic B-spline basis, Bspline.h:253-276
ORC_HOT void basis(real u, real* b) {
    const real inv6 = real(1.) / real(6.);
    const real u2 = u * u, t = 1 - u;
    b[0] = inv6 * t * t * t;
    b[1] = inv6 * (3 * u2 * (u - 2) + 4);
    b[2] = inv6 * (3 * u * (1 + u - u2) + 1);
    b[3] = inv6 * u2 * u;
}
ORC_HOT void dbasis(real u, real* d) {
    const real u2 = u * u, t = 1 - u;
    d[0] = -real(0.5) * t * t;
    d[1] = real(0.5) * (3 * u2 - 4 * u);
    d[2] = real(0.5) * (-3 * u2 + 2 * u + 1);
    d[3] = real(0.5) * u2;
}
ORC_HOT int modulo(int i, int m) {  // Bspline.h:83-89
    int r = i % m;
    if (r < 0) r += m;
    return r;
}
// periodic tap offsets: choose_inds_pbc over make_inds_pbc tables, Bspline.h:39-60
ORC_HOT void tap_offsets(int i, int n, int* o) {
    if (i >= n - 3) {
        const int ii = i + 4 - n;
        for (int k = 0; k < 4; k++) o[k] = (k < 4 - ii) ? k : k - n;
    } else {
        for (int k = 0; k < 4; k++) o[k] = k;
    }
}

// getForce_Bspline + fe3d_pbc_comb3 (GridFF.h:355-402, Bspline.h:742-809)
ORC_HOT real grid_sample(const Sys& S, V p, const real* PLQ, V& fout) {
    const int nx = S.gn[0], ny = S.gn[1], nz = S.gn[2];
    const real ux = (p.x - S.gp0[0]) / S.gstep[0];
    const real uy = (p.y - S.gp0[1]) / S.gstep[1];
    const real uz = (p.z - S.gp0[2]) / S.gstep[2];
    int ix = (int)ux, iy = (int)uy, iz = (int)uz;
    if (ux < 0) ix--;
    if (uy < 0) iy--;
    const real tx = ux - ix, ty = uy - iy, tz = uz - iz;
    iz = std::min(std::max(iz, 3), nz - 5);
    ix = modulo(ix - 1, nx);
    iy = modulo(iy - 1, ny);
    int qx[4], qy[4];
    tap_offsets(ix, nx, qx);
    tap_offsets(iy, ny, qy);
    real bz[4], dz[4], by[4], dy[4], bx[4], dx[4];
    basis(tz, bz);
    dbasis(tz, dz);
    basis(ty, by);
    dbasis(ty, dy);
    basis(tx, bx);
    dbasis(tx, dx);
    const int i0 = (iz - 1) + nz * (iy + ny * ix);
    real Ex[4][3];  // per x tap: (Fy, Fz, E)  (fe2d_comb3)
    for (int a = 0; a < 4; a++) {
        real fe[4][2];
        for (int b = 0; b < 4; b++) {
            const real* E = S.gc + 3 * (size_t)(i0 + qx[a] * ny * nz + qy[b] * nz);
            real cs[4];
            for (int c = 0; c < 4; c++)
                cs[c] = PLQ[0] * E[c * 3] + PLQ[1] * E[c * 3 + 1] + PLQ[2] * E[c * 3 + 2];
            fe[b][0] = bz[0] * cs[0] + bz[1] * cs[1] + bz[2] * cs[2] + bz[3] * cs[3];
            fe[b][1] = dz[0] * cs[0] + dz[1] * cs[1] + dz[2] * cs[2] + dz[3] * cs[3];
        }
        Ex[a][0] = fe[0][0] * dy[0] + fe[1][0] * dy[1] + fe[2][0] * dy[2] + fe[3][0] * dy[3];
        Ex[a][1] = fe[0][1] * by[0] + fe[1][1] * by[1] + fe[2][1] * by[2] + fe[3][1] * by[3];
        Ex[a][2] = fe[0][0] * by[0] + fe[1][0] * by[1] + fe[2][0] * by[2] + fe[3][0] * by[3];
    }
    const real Fx = dx[0] * Ex[0][2] + dx[1] * Ex[1][2] + dx[2] * Ex[2][2] + dx[3] * Ex[3][2];
    const real Fy = bx[0] * Ex[0][0] + bx[1] * Ex[1][0] + bx[2] * Ex[2][0] + bx[3] * Ex[3][0];
    const real Fz = bx[0] * Ex[0][1] + bx[1] * Ex[1][1] + bx[2] * Ex[2][1] + bx[3] * Ex[3][1];
    const real E = bx[0] * Ex[0][2] + bx[1] * Ex[1][2] + bx[2] * Ex[2][2] + bx[3] * Ex[3][2];
    fout += V{-Fx / S.gstep[0], -Fy / S.gstep[1], -Fz / S.gstep[2]};
    return E;
}

// REQ2PLQ_d, Forces.h:611-619
void req2plq(const real* REQ, real alpha, real* PLQ) {
    real e = std::exp(alpha * REQ[0]);
    real cL = e * REQ[1];
    PLQ[0] = e * cL;
    PLQ[1] = cL;
    PLQ[2] = REQ[2];
    PLQ[3] = e * e * REQ[3];
}

// Substrate atoms, GridFF::evalMorsePBC (GridFF.h:1229-1261) with
// addAtomicForceMorseQ (Forces.h:540-561) and combineREQ (Forces.h:263-268);
// the pair's Morse amplitude E0*e is rounded to float as the reference does
// (`float Ae`).  Adds the force to fapos[ia], returns the energy.
real surf_atom(Sys& S, int ia) {
    const V pi = ld(&S.pos[ia * 3]);
    const real* Ri = S.REQ + ia * 4;
    const real K = g_surf.K;
    real E = 0;
    V f{0, 0, 0};
    for (int j = 0; j < g_surf.natoms; j++) {
        const real* Rj = &g_surf.REQs[j * 4];
        const real r0 = Rj[0] + Ri[0], E0 = Rj[1] * Ri[1], qq = Rj[2] * Ri[2];
        const V dp0 = pi - ld(&g_surf.apos[j * 3]);
        for (int ipbc = 0; ipbc < g_surf.npbc; ipbc++) {
            const V dp = dp0 + ld(&g_surf.shifts[ipbc * 3]);
            const real r2 = dot(dp, dp);
            const real ir2 = 1 / r2;
            const real r = std::sqrt(r2 + real(1e-32));
            const real ir = std::sqrt(ir2);
            const real e = std::exp(K * (r - r0));
            const float Ae = (float)(E0 * e);
            const real fM = Ae * 2 * K * (e - 1);
            const real EM = Ae * (e - 2);
            const real Eel = COULOMB_K * qq * ir;
            f -= dp * (fM / r - Eel * ir2);
            E += EM + Eel;
        }
    }
    st(&S.f[ia * 3], ld(&S.f[ia * 3]) + f);
    return E;
}

// eval_no_omp (MolWorld_sp3.h:2178-2245) + assemble_atom (MMFFsp3_loc.h:998-1021)
real eval_sys(Sys& S) {
    const int N = S.N, M = S.M;
    S.f.assign((size_t)(N + M) * 3, real(0.0));
    S.fng.assign((size_t)M * 12, real(0.0));
    S.fngpi.assign((size_t)M * 12, real(0.0));
    real E = 0;
    for (int ia = 0; ia < N; ia++) {
        st(&S.f[ia * 3], V{0, 0, 0});
        if (ia < M) st(&S.f[(N + ia) * 3], V{0, 0, 0});
        if (ia < M) E += eval_node(S, ia);
        if (S.bNonBonded) E += nb_atom(S, ia);
        if (S.gc) {
            real plq[4];
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
            const real* c = S.constr + ia * 4;
            const real* k = S.constrK + ia * 4;
            const V p = ld(&S.pos[ia * 3]);
            const V fc{(c[0] - p.x) * std::max(k[0], real(0)), (c[1] - p.y) * std::max(k[1], real(0)),
                       (c[2] - p.z) * std::max(k[2], real(0))};
            st(&S.f[ia * 3], ld(&S.f[ia * 3]) + fc);
        }
        for (int d = 0; d < 3; d++) {
            const real kw = S.bbox[6 + d];
            if (!(kw > 0)) continue;
            const real x = S.pos[ia * 3 + d], lo = S.bbox[d], hi = S.bbox[3 + d];
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

void setup(Sys& S, int natoms, int nnode, const real* apos, const real* REQs, const int* neighs,
           const real* apars, const real* bLs, const real* bKs, const real* Ksp, const real* Kpp,
           const int* sw, int gnx, int gny, int gnz, const real* gpos0, const real* gstep, const real* gcoefs,
           real alpha) {
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
    real dt, dt_max, dt_min, damping, damp_max, finc, fdec, falpha;
    int minLastNeg, lastNeg;
    real cv, cf, renorm;
};

// FIRE_update_params (DynamicOpt.cpp:250-267), damp_func_FIRE (:210-222)
void fire_update(Fire& F, real vf, real vv, real ff) {
    const real f_len = std::sqrt(ff), v_len = std::sqrt(vv);
    const real cos_vf = vf / (f_len * v_len + real(1e-32));
    F.renorm = v_len / (f_len + real(1e-32));
    real cf;
    if (cos_vf < real(0.0)) {
        F.cv = real(0.0);  // cv_kill
        cf = real(0.0);
    } else {
        F.cv = 1 - F.damping;
        cf = F.damping;
    }
    F.cf = F.renorm * cf;
    if (cos_vf < real(0.0)) {
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
int relax_sys(Sys& S, const real* fire, real Fconv, int nmax, int conv_mode, real* E_out, real* traj,
              int ntraj, real* fmax_out) {
    const real Flimit = fire[7];
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
    std::vector<real> vel((size_t)V_ * 3, real(0.0));
    const real F2conv = Fconv * Fconv;
    int itr = 0;
    real E = 0, fm2 = 0;
    for (itr = 0; itr < nmax; itr++) {
        E = eval_sys(S);
        real vf = 0, vv = 0, ff = 0;
        fm2 = 0;
        for (int i = 0; i < V_; i++) fm2 = std::max(fm2, dot(ld(&S.f[i * 3]), ld(&S.f[i * 3])));
        if (Flimit > 0) {  // per-vec clamp of the GPU update kernel, relax_multi.cl:980-983
            for (int i = 0; i < V_; i++) {
                V f = ld(&S.f[i * 3]);
                real f2 = dot(f, f);
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
        const real dt = (real)(float)F.dt;
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
            if (bPi) p = p * (real(1.0) / std::sqrt(dot(p, p)));
            st(&vel[i * 3], v);
            st(&S.pos[i * 3], p);
        }
        if (traj && itr < ntraj) std::memcpy(traj + (size_t)itr * V_ * 3, S.pos.data(), sizeof(real) * V_ * 3);
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
real drive_xi(uint32_t seed, uint32_t sys, uint32_t vec, uint32_t axis, uint32_t step) {
    uint32_t h = fmix32(seed + 0x9E3779B9u);
    h = fmix32(h ^ sys);
    h = fmix32(h ^ (vec * 3u + axis));
    h = fmix32(h ^ step) >> 8;
    return ((real)h * (real(2.0) / real(16777216.0)) - real(1.0)) * std::sqrt(real(3.0));
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
int explore_sys(Sys& S, const real* fire, real Fconv, int nmax, real T, real gamma, int nExplore, int nRelax,
                uint32_t seed, uint32_t sysidx, int kmax, real* minE, real* minpos, real* traj, int ntraj) {
    const real kB = real(8.617333262145e-5);  // eV/K, cpp/common/constants.h:12
    const real Flimit = fire[7];
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
    std::vector<real> vel((size_t)V_ * 3, real(0.0));
    const real F2conv = Fconv * Fconv;
    int phase = 0, pstep = 0, nfound = 0;
    for (int step = 0; step <= nmax; step++) {
        const real E = eval_sys(S);
        real fm2 = 0, vf = 0, vv = 0, ff = 0;
        for (int i = 0; i < V_; i++) fm2 = std::max(fm2, dot(ld(&S.f[i * 3]), ld(&S.f[i * 3])));
        if (Flimit > 0)
            for (int i = 0; i < V_; i++) {
                V f = ld(&S.f[i * 3]);
                real f2 = dot(f, f);
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
                        if (minpos) std::memcpy(minpos + (size_t)nfound * V_ * 3, S.pos.data(), sizeof(real) * V_ * 3);
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
        real dt, cv, cfe;
        if (drive) {
            dt = (real)(float)F.dt_max;
            cv = real(1.0);
            cfe = real(0.0);
        } else {
            if (reset) fire_reset();
            fire_update(F, vf, vv, ff);
            dt = (real)(float)F.dt;
            cv = F.cv;
            cfe = F.renorm * F.cf;
        }
        const real amp = drive ? std::sqrt(real(2.0) * kB * std::max(T, real(0)) * std::max(gamma, real(0)) / dt) : real(0.0);
        for (int i = 0; i < V_; i++) {
            V f = ld(&S.f[i * 3]);
            V v = ld(&vel[i * 3]);
            V p = ld(&S.pos[i * 3]);
            const bool bPi = i >= N;
            if (bPi) {
                f -= p * dot(p, f);
            } else if (drive) {
                f -= v * std::max(gamma, real(0));
                f += V{amp * drive_xi(seed, sysidx, (uint32_t)i, 0u, (uint32_t)step),
                       amp * drive_xi(seed, sysidx, (uint32_t)i, 1u, (uint32_t)step),
                       amp * drive_xi(seed, sysidx, (uint32_t)i, 2u, (uint32_t)step)};
            }
            v = v * cv;
            v += f * (dt + cfe);
            if (bPi) v -= p * dot(p, v);
            p += v * dt;
            if (bPi) p = p * (real(1.0) / std::sqrt(dot(p, p)));
            st(&vel[i * 3], v);
            st(&S.pos[i * 3], p);
        }
        pstep++;
        if (traj && step < ntraj) std::memcpy(traj + (size_t)step * V_ * 3, S.pos.data(), sizeof(real) * V_ * 3);
    }
    return nfound;
}

// Batch of CSR-packed systems on `nthreads` host threads, one system per task,
// dynamic schedule (the structure of MolWorld_sp3_multi::run_omp_ocl,
// :1948-2027).  apos is updated in place; steps/E/Fmax out.
void relax_batch(int nsys, const int* natoms, const int* nnode, real* apos, const real* REQs, const int* neighs,
                 const real* apars, const real* bLs, const real* bKs, const real* Ksp, const real* Kpp, const int* sw,
                 int gnx, int gny, int gnz, const real* gpos0, const real* gstep, const real* gcoefs, real alpha,
                 const real* fire, real Fconv, int nmax, int conv_mode, int nthreads, int* steps_out, real* E_out,
                 real* Fmax_out) {
    std::vector<size_t> a0(nsys + 1, 0), n0(nsys + 1, 0), v0(nsys + 1, 0);
    for (int s = 0; s < nsys; s++) {
        a0[s + 1] = a0[s] + natoms[s];
        n0[s + 1] = n0[s] + nnode[s];
        v0[s + 1] = v0[s] + natoms[s] + nnode[s];
    }
    std::atomic<int> next{0};
    auto worker = [&]() {
        // single precision: flush denormals (FTZ | DAZ), as FP32 GPU code does;
        // underflowing long-range terms otherwise take the slow denormal path
        if (sizeof(real) == 4) _mm_setcsr(_mm_getcsr() | 0x8040);
        for (;;) {
            int s = next.fetch_add(1);
            if (s >= nsys) break;
            Sys S;
            setup(S, natoms[s], nnode[s], apos + v0[s] * 3, REQs + a0[s] * 4, neighs + a0[s] * 4, apars + n0[s] * 4,
                  bLs + n0[s] * 4, bKs + n0[s] * 4, Ksp + n0[s] * 4, Kpp + n0[s] * 4, sw, gnx, gny, gnz, gpos0, gstep,
                  gcoefs, alpha);
            real E = 0, fm = 0;
            int n = relax_sys(S, fire, Fconv, nmax, conv_mode, &E, nullptr, 0, &fm);
            std::memcpy(apos + v0[s] * 3, S.pos.data(), sizeof(real) * S.pos.size());
            if (steps_out) steps_out[s] = n;
            if (E_out) E_out[s] = E;
            if (Fmax_out) Fmax_out[s] = fm;
        }
    };
    nthreads = std::max(1, nthreads);
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(worker);
    for (auto& t : th) t.join();
}
