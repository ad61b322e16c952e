// FP32 device force-field terms of the multi-system MMFF relaxation.
//
// Each function restates one term of the reference CPU force field in single
// precision for gfx950; the reference line ranges are given per function.
// The same functions serve the split kernels (one launch per phase) and the
// persistent relaxation kernel (all phases in one workgroup).
#pragma once
#include <hip/hip_runtime.h>

namespace fcd {

constexpr float COULOMB_K = 14.3996448915f;  // eV*A/e^2 (cpp/common/constants.h:9)

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 mk(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 xyz(const float4& a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ void operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
__device__ __forceinline__ void operator-=(V3& a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 zero3() { return V3{0.f, 0.f, 0.f}; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Wave shift by one lane (DPP wave_shl:1 with bound_ctrl): lane l receives
// lane l+1's value, lane 63 receives 0.  Needs all 64 lanes active.
__device__ __forceinline__ float dpp_shl1_zero(float src) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, src), 0x130, 0xF, 0xF,
                                                                 true));
}

// Switches of the bonded terms (setSwitches_multi, MMFFmulti_lib.cpp:316-329).
struct FFSwitch {
    int doBonds, doPiPiI, doPiSigma, doAngles;
};

// -------------------------------------------------------------------------
// Bond stretch, Forces.h:137-143 (evalBond): f = h*(dl*k), E = k dl^2 / 2.
// pi-pi alignment, Forces.h:230-241 (evalPiAling) with ir1=ir2=1.
__device__ __forceinline__ float pi_align(V3 h1, V3 h2, float K, V3& f1, V3& f2) {
    float c = dot(h1, h2);
    f1 = h2 - h1 * c;
    f2 = h1 - h2 * c;
    bool sgn = c < 0.f;
    if (sgn) c = -c;
    float E = K * (1.f - c);
    float fang = sgn ? -K : K;
    f1 = f1 * fang;
    f2 = f2 * fang;
    return E;
}
// pi-sigma orthogonality, Forces.h:218-228 (evalAngleCos).
__device__ __forceinline__ float angle_cos(V3 h1, V3 h2, float ir1, float ir2, float K, float c0, V3& f1, V3& f2) {
    float c = dot(h1, h2);
    f1 = h2 - h1 * c;
    f2 = h1 - h2 * c;
    float c_ = c - c0;
    float E = K * c_ * c_;
    float fang = -K * c_ * 2.f;
    f1 = f1 * (fang * ir1);
    f2 = f2 * (fang * ir2);
    return E;
}
// sigma-sigma angle in the cos(theta/2) form, Forces.h:190-216
// (evalAngleCosHalf); cs0 = (cos(a0/2), sin(a0/2)).
__device__ __forceinline__ float angle_cos_half(V3 h1, V3 h2, float ir1, float ir2, float cs0x, float cs0y, float k,
                                                V3& f1, V3& f2) {
    V3 h = h1 + h2;
    float c2 = dot(h, h) * 0.25f;
    float s2 = 1.f - c2 + 1e-14f;
    float c = sqrtf(c2);
    float s = sqrtf(s2);
    // cs = cs0 udiv (c,s)  (Vec2.h:130)
    float csx = cs0x * c + cs0y * s;
    float csy = cs0y * c - cs0x * s;
    // k*(1 - csx) written as k*csy^2/(1 + csx) (|cs| = 1): no cancellation
    // near the equilibrium angle, where FP32 1 - csx loses the energy
    float E = k * csy * csy / (1.f + csx);
    float fr = -k * csy;
    c2 *= -2.f;
    fr /= 4.f * c * s;
    float fr1 = fr * ir1;
    float fr2 = fr * ir2;
    f1 = h * fr1 + h1 * (fr1 * c2);
    f2 = h * fr2 + h2 * (fr2 * c2);
    return E;
}

// -------------------------------------------------------------------------
// Bonded forces of one node, MMFFsp3_loc::eval_atom (MMFFsp3_loc.h:250-575)
// in the bNonBondNeighs strategy (bSubtractBondNonBond=false, ForceField.h:
// 168-184).  The 1-3 non-bonded subtraction of that function is not done
// here: the non-bonded sum excludes those pairs directly (see
// make_exclusions), which gives the same total force and energy.
// Outputs: recoils on the 4 neighbours (fbs) and on their pi-vectors (fps),
// the node's own force (fa) and its pi force (fpi).  Returns the energy.
// `pos(i)` returns the position of vec i (atom or, for i>=natoms, pi-vector);
// `nshift(k)` the periodic image shift of neighbour slot k (pbc_shifts of its
// neighCell, MMFFsp3_loc.h:325-329; zero in vacuum).
template <class PosFn, class ShiftFn>
__device__ __forceinline__ float eval_node(int ia, int natoms, int nnode, PosFn pos, ShiftFn nshift, const int4 ng,
                                           const float4 apar, const float4 bL, const float4 bK, const float4 ksp,
                                           const float4 kpp, const FFSwitch sw, V3* fbs, V3* fps, V3& fa, V3& fpi) {
    const V3 pa = pos(ia);
    const V3 hpi = pos(natoms + ia);
    const int ings[4] = {ng.x, ng.y, ng.z, ng.w};
    const float bLs[4] = {bL.x, bL.y, bL.z, bL.w};
    const float bKs[4] = {bK.x, bK.y, bK.z, bK.w};
    const float Ksp[4] = {ksp.x, ksp.y, ksp.z, ksp.w};
    const float Kpp[4] = {kpp.x, kpp.y, kpp.z, kpp.w};
    const float piC0 = apar.w;
    float E = 0.f;
    fa = zero3();
    fpi = zero3();
    V3 hs[4];
    float ir[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        fbs[i] = zero3();
        fps[i] = zero3();
    }
    int nb = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int ing = ings[i];
        if (ing < 0) break;
        nb = i + 1;
        V3 h = pos(ing) + nshift(i) - pa;
        const float l = sqrtf(dot(h, h));
        const float il = 1.f / l;
        h = h * il;
        hs[i] = h;
        ir[i] = il;
        V3 f1, f2;
        if (ia < ing) {
            if (sw.doBonds) {
                const float dl = l - bLs[i];
                const float fr = dl * bKs[i];
                f1 = h * fr;
                E += fr * dl * 0.5f;
                fbs[i] -= f1;
                fa += f1;
            }
            const float k = Kpp[i];
            if (sw.doPiPiI && (ing < nnode) && (k > 1e-6f)) {
                E += pi_align(hpi, pos(natoms + ing), k, f1, f2);
                fpi += f1;
                fps[i] += f2;
            }
        }
        const float k = Ksp[i];
        if (sw.doPiSigma && (k > 1e-6f)) {
            E += angle_cos(hpi, h, 1.f, il, k, piC0, f1, f2);
            fpi += f1;
            fa -= f2;
            fbs[i] += f2;
        }
    }
    if (sw.doAngles) {
        const float ssK = apar.z;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (i + 1 >= nb) break;
#pragma unroll
            for (int j = i + 1; j < 4; j++) {
                if (j >= nb) break;
                V3 f1, f2;
                E += angle_cos_half(hs[i], hs[j], ir[i], ir[j], apar.x, apar.y, ssK, f1, f2);
                fa -= f1 + f2;
                fbs[i] += f1;
                fbs[j] += f2;
            }
        }
    }
    return E;
}
// Vacuum molecules: no image shifts.
template <class PosFn>
__device__ __forceinline__ float eval_node(int ia, int natoms, int nnode, PosFn pos, const int4 ng, const float4 apar,
                                           const float4 bL, const float4 bK, const float4 ksp, const float4 kpp,
                                           const FFSwitch sw, V3* fbs, V3* fps, V3& fa, V3& fpi) {
    return eval_node(ia, natoms, nnode, pos, [](int) { return zero3(); }, ng, apar, bL, bK, ksp, kpp, sw, fbs, fps, fa,
                     fpi);
}

// -------------------------------------------------------------------------
// Directed non-bonded pair, Forces.h:341-360 (getLJQH): bare Coulomb, LJ and
// the H-bond correction; REQ mixing Forces.h:270 (_mixREQ).  d = pj - pi;
// returns the force on i and the (unhalved) pair energy.
struct Req {
    float R, E, kQ, H;  // R0, sqrt(E0), COULOMB_K*Q, H
};
// 1/r from one v_rsq_f32 (1 ulp); the H-bond term is compiled in only for
// batches that carry H-bond charges (HB=true).
// PRECISE adds one Newton step to 1/r (energy passes: the Coulomb sums cancel
// to a small fraction of their terms, so a biased 1-ulp 1/r shows up there).
__device__ __forceinline__ float rsq_nr(float x) {
    const float y = __builtin_amdgcn_rsqf(x);
    return y * (1.5f - 0.5f * x * y * y);
}
template <bool HB = true, bool PRECISE = false>
__device__ __forceinline__ void lj_pair(V3 d, float Rij, float Eij, float kQQ, float Hij, float& fr, float& e) {
    const float r2 = dot(d, d);
    const float ir = PRECISE ? rsq_nr(r2) : __builtin_amdgcn_rsqf(r2);
    const float ir2 = ir * ir;
    const float Eq = kQQ * ir;
    const float u2 = (Rij * Rij) * ir2;
    const float u6 = u2 * u2 * u2;
    const float vdW = u6 * Eij;
    if (HB) {
        const float H = u6 * u6 * ((Hij < 0.f) ? Hij * Eij : 0.f);
        e = Eq + (u6 - 2.f) * vdW + H;
        fr = (Eq + ((u6 - 1.f) * vdW + H) * 12.f) * ir2;  // force on i = d * (-fr)
    } else {
        e = Eq + (u6 - 2.f) * vdW;
        fr = (Eq + (u6 - 1.f) * vdW * 12.f) * ir2;
    }
}

// -------------------------------------------------------------------------
// Cubic B-spline GridFF sample, GridFF::getForce_Bspline (GridFF.h:355-402)
// -> Bspline::fe3d_pbc_comb3 (Bspline.h:742-809), basis/dbasis (:253-276).
// x,y periodic; z index truncated toward zero and clamped to [3, nz-5] while
// tz keeps the unclamped index (Bspline.h:746-759).
struct GridDev {
    const float4* coefs;  // (nx,ny,nz) float4 (P, L, Q, 0), z fastest
    int nx, ny, nz;
    float p0x, p0y, p0z;
    float isx, isy, isz;  // 1/step
};

__device__ __forceinline__ void bs_basis(float u, float* b) {
    const float inv6 = 1.f / 6.f;
    const float u2 = u * u;
    const float t = 1.f - u;
    b[0] = inv6 * t * t * t;
    b[1] = inv6 * (3.f * u2 * (u - 2.f) + 4.f);
    b[2] = inv6 * (3.f * u * (1.f + u - u2) + 1.f);
    b[3] = inv6 * u2 * u;
}
__device__ __forceinline__ void bs_dbasis(float u, float* d) {
    const float u2 = u * u;
    const float t = 1.f - u;
    d[0] = -0.5f * t * t;
    d[1] = 0.5f * (3.f * u2 - 4.f * u);
    d[2] = 0.5f * (-3.f * u2 + 2.f * u + 1.f);
    d[3] = 0.5f * u2;
}
__device__ __forceinline__ int imod(int i, int m) {
    int r = i % m;
    return r < 0 ? r + m : r;
}

// Returns E; adds the force to f.  plq = (cP, cL, cQ) of the atom.
__device__ __forceinline__ float grid_sample(const GridDev& g, V3 p, V3 plq, V3& f) {
    const float ux = (p.x - g.p0x) * g.isx;
    const float uy = (p.y - g.p0y) * g.isy;
    const float uz = (p.z - g.p0z) * g.isz;
    int ix = (int)ux, iy = (int)uy, iz = (int)uz;
    if (ux < 0.f) ix--;
    if (uy < 0.f) iy--;
    const float tx = ux - ix, ty = uy - iy, tz = uz - iz;
    iz = min(max(iz, 3), g.nz - 5);
    ix = imod(ix - 1, g.nx);
    iy = imod(iy - 1, g.ny);
    float bx[4], dx[4], by[4], dy[4], bz[4], dz[4];
    bs_basis(tx, bx);
    bs_dbasis(tx, dx);
    bs_basis(ty, by);
    bs_dbasis(ty, dy);
    bs_basis(tz, bz);
    bs_dbasis(tz, dz);
    float Fx = 0.f, Fy = 0.f, Fz = 0.f, E = 0.f;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        int xa = ix + a;
        if (xa >= g.nx) xa -= g.nx;
        float ey = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            int yb = iy + b;
            if (yb >= g.ny) yb -= g.ny;
            const float4* col = g.coefs + ((size_t)(xa * g.ny + yb) * g.nz + (iz - 1));
            float e = 0.f, de = 0.f;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const float4 q = col[c];
                const float cs = plq.x * q.x + plq.y * q.y + plq.z * q.z;
                e += bz[c] * cs;
                de += dz[c] * cs;
            }
            fy += e * dy[b];
            fz += de * by[b];
            ey += e * by[b];
        }
        Fx += ey * dx[a];
        Fy += fy * bx[a];
        Fz += fz * bx[a];
        E += ey * bx[a];
    }
    f.x -= Fx * g.isx;
    f.y -= Fy * g.isy;
    f.z -= Fz * g.isz;
    return E;
}

// The same sample with the atom's 4x4x4 window of channel-combined
// coefficients (cs = plq . c, 64 floats = 256 B) kept in a per-atom cache
// slot in global memory.  While the atom stays in the same grid cell (the
// window base `key` is unchanged) the 64 gathers from the 10 MB grid become
// one contiguous 256 B read that stays in L2; on a cell change the window is
// gathered, combined and written back.  Same arithmetic as grid_sample.
// Cubic B-spline basis value and derivative of index a (0..3) at u: the
// slab loop below runs rolled, so the four weights are computed per slab.
__device__ __forceinline__ void bs_basis_a(float u, int a, float& b, float& d) {
    const float inv6 = 1.f / 6.f;
    const float u2 = u * u;
    const float t = 1.f - u;
    switch (a) {
        case 0: b = inv6 * t * t * t; d = -0.5f * t * t; break;
        case 1: b = inv6 * (3.f * u2 * (u - 2.f) + 4.f); d = 0.5f * (3.f * u2 - 4.f * u); break;
        case 2: b = inv6 * (3.f * u * (1.f + u - u2) + 1.f); d = 0.5f * (-3.f * u2 + 2.f * u + 1.f); break;
        default: b = inv6 * u2 * u; d = 0.5f * u2; break;
    }
}

__device__ __forceinline__ float grid_sample_cached(const GridDev& g, V3 p, V3 plq, float4* cache, int stride, int& key,
                                                    V3& f) {
    const float ux = (p.x - g.p0x) * g.isx;
    const float uy = (p.y - g.p0y) * g.isy;
    const float uz = (p.z - g.p0z) * g.isz;
    int ix = (int)ux, iy = (int)uy, iz = (int)uz;
    if (ux < 0.f) ix--;
    if (uy < 0.f) iy--;
    const float tx = ux - ix, ty = uy - iy, tz = uz - iz;
    iz = min(max(iz, 3), g.nz - 5);
    ix = imod(ix - 1, g.nx);
    iy = imod(iy - 1, g.ny);
    const int k = (ix * g.ny + iy) * g.nz + iz;
    const bool hit = (k == key);
    key = k;
    float by[4], dy[4], bz[4], dz[4];
    bs_basis(ty, by);
    bs_dbasis(ty, dy);
    bs_basis(tz, bz);
    bs_dbasis(tz, dz);
    float Fx = 0.f, Fy = 0.f, Fz = 0.f, E = 0.f;
#pragma unroll 1
    for (int a = 0; a < 4; a++) {  // x-slabs one at a time: 16 coefficients live
        float cs[16];
        if (hit) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float4 v = cache[(a * 4 + q) * stride];
                cs[q * 4 + 0] = v.x;
                cs[q * 4 + 1] = v.y;
                cs[q * 4 + 2] = v.z;
                cs[q * 4 + 3] = v.w;
            }
        } else {
            int xa = ix + a;
            if (xa >= g.nx) xa -= g.nx;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                int yb = iy + b;
                if (yb >= g.ny) yb -= g.ny;
                const float4* col = g.coefs + ((size_t)(xa * g.ny + yb) * g.nz + (iz - 1));
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const float4 q = col[c];
                    cs[b * 4 + c] = plq.x * q.x + plq.y * q.y + plq.z * q.z;
                }
            }
#pragma unroll
            for (int q = 0; q < 4; q++)
                cache[(a * 4 + q) * stride] = make_float4(cs[q * 4], cs[q * 4 + 1], cs[q * 4 + 2], cs[q * 4 + 3]);
        }
        float ey = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            float e = 0.f, de = 0.f;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                e += bz[c] * cs[b * 4 + c];
                de += dz[c] * cs[b * 4 + c];
            }
            fy += e * dy[b];
            fz += de * by[b];
            ey += e * by[b];
        }
        float bxa, dxa;
        bs_basis_a(tx, a, bxa, dxa);
        Fx += ey * dxa;
        Fy += fy * bxa;
        Fz += fz * bxa;
        E += ey * bxa;
    }
    f.x -= Fx * g.isx;
    f.y -= Fy * g.isy;
    f.z -= Fz * g.isz;
    return E;
}

// -------------------------------------------------------------------------
// Positional restraint and bounding wall of one atom at p, the terms of the
// reference's update kernel (updateAtomsMMFFf4, relax_multi.cl:997-1009):
//   restraint (c.w > 0):  f += (c.xyz - p) * max(k.xyz, 0)
//   wall on axis d with stiffness bc_d > 0:  f_d += (lo_d - p_d) bc_d below
//   lo_d, (hi_d - p_d) bc_d above hi_d  (the reference enables z only; x and
//   y follow the same rule when their stiffness is positive).
__device__ __forceinline__ V3 restraint_force(V3 p, float4 c, float4 k) {
    if (!(c.w > 0.f)) return zero3();
    return mk((c.x - p.x) * fmaxf(k.x, 0.f), (c.y - p.y) * fmaxf(k.y, 0.f), (c.z - p.z) * fmaxf(k.z, 0.f));
}
__device__ __forceinline__ float wall1(float x, float lo, float hi, float k) {
    if (!(k > 0.f)) return 0.f;
    return (x < lo) ? (lo - x) * k : ((x > hi) ? (hi - x) * k : 0.f);
}
__device__ __forceinline__ V3 wall_force(V3 p, float4 lo, float4 hi, float4 k) {
    return mk(wall1(p.x, lo.x, hi.x, k.x), wall1(p.y, lo.y, hi.y, k.y), wall1(p.z, lo.z, hi.z, k.z));
}

// -------------------------------------------------------------------------
// FIRE parameters (DynamicOpt.h:47-95 defaults; set_opt, MMFFmulti_lib.cpp:141)
struct FireParams {
    float dt_max, dt_min, damp_max, finc, fdec, falpha, cv_kill;
    int minLastNeg;
};
// Per-system FIRE state.
struct FireState {
    float dt, damping;
    int lastNeg;
};
// DynamicOpt::FIRE_update_params (DynamicOpt.cpp:250-267) with
// damp_func_FIRE (:210-222).  Returns (cv, cf_eff) where the velocity update
// is v = cv*v + (dt + cf_eff)*f and cf_eff = renorm*cf as passed by
// MolWorld_sp3::run_no_omp (MolWorld_sp3.h:2426).
__device__ __forceinline__ void fire_update(const FireParams& P, FireState& S, float vf, float vv, float ff, float& cv,
                                            float& cfe) {
    const float f_len = sqrtf(ff);
    const float v_len = sqrtf(vv);
    const float cos_vf = vf / (f_len * v_len + 1e-32f);
    const float renorm = v_len / (f_len + 1e-32f);
    float cf;
    if (cos_vf < 0.f) {
        cv = P.cv_kill;
        cf = 0.f;
    } else {
        cv = 1.f - S.damping;
        cf = S.damping;
    }
    cf *= renorm;
    if (cos_vf < 0.f) {
        S.dt = fmaxf(S.dt * P.fdec, P.dt_min);
        S.lastNeg = 0;
        S.damping = P.damp_max;
    } else {
        if (S.lastNeg > P.minLastNeg) {
            S.dt = fminf(S.dt * P.finc, P.dt_max);
            S.damping = S.damping * P.falpha;
        }
        S.lastNeg++;
    }
    cfe = renorm * cf;
}

// -------------------------------------------------------------------------
// Langevin drive of the exploring phase (updateAtomsMMFFf4,
// relax_multi.cl:1160-1178): f += -gamma*v + xi*sqrt(2*kB*T*gamma/dt) on atoms
// (pi-vectors are not driven), with xi a unit-variance random number per
// component.  xi comes from a counter-based hash of (seed, system, vec, axis,
// step) -- murmur3's 32-bit finaliser applied after each field -- mapped to a
// uniform on [-sqrt(3), sqrt(3)); integer arithmetic only, so the FP64 oracle
// draws the same numbers.
constexpr float BOLTZMANN_K = 8.617333262145e-5f;  // eV/K (cpp/common/constants.h:12)
__host__ __device__ inline unsigned fmix32(unsigned h) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}
__host__ __device__ inline unsigned drive_hash(unsigned seed, unsigned sys, unsigned vec, unsigned axis,
                                               unsigned step) {
    unsigned h = fmix32(seed + 0x9E3779B9u);
    h = fmix32(h ^ sys);
    h = fmix32(h ^ (vec * 3u + axis));
    return fmix32(h ^ step);
}
__device__ __forceinline__ float drive_xi(unsigned seed, unsigned sys, unsigned vec, unsigned axis, unsigned step) {
    const unsigned h = drive_hash(seed, sys, vec, axis, step) >> 8;  // 24 bits
    return ((float)h * (2.f / 16777216.f) - 1.f) * 1.7320508f;
}

}  // namespace fcd
